"""The replicate batch at large sizes.

- At its largest supported size: 14 runs of the 50M-agent workload take
  700,055,552 of the 715,827,882 agent slots a batch holds (each run rounded
  up to 8192-slot tiles; include/pandemic/pt_batch.h).  Over 3 weekdays the
  first and last replicates equal their single runs day by day, and a 15th
  replicate is refused with an error rather than run.
- As the surrogate sweep re-simulates: two intervention scenarios at 50M
  agents (initial infected per million), two replicates each, in one batch,
  equal their single runs over a week, weekend included."""
import numpy as np
import pytest

from pandemicthrust_amd import PandemicBatch, PandemicSim, Params
from pandemicthrust_amd import sweep

pytestmark = pytest.mark.gpu

ROW = ["3", "1.8", "1.3", "10", "10", "21.24", "21.24", "1", "0"]   # 50,002,479 agents, days 0-2


def _slots(n_people):
    return int(sum(-(-int(x) // sweep.SLOT_TILE) * sweep.SLOT_TILE for x in n_people))


def test_batch_at_the_slot_limit_matches_single_runs():
    p = Params.from_row(ROW)
    seeds = [101 + i for i in range(14)]
    print("batch: 14 x 50M agents, 3 days", flush=True)
    b = PandemicBatch([p] * len(seeds), seeds)
    n_people = b.setup()
    slots = _slots(n_people)
    assert b.info()["slots"] == slots
    assert sweep.MAX_SLOTS - slots < slots // len(seeds)        # no room for one more run
    res = b.run()
    b.close()
    for i in (0, len(seeds) - 1):
        print("single run, seed %d" % seeds[i], flush=True)
        s = PandemicSim(p, seeds[i])
        s.setup_sim()
        counts, inf = s.run()
        s.close()
        assert len(inf) == 3
        assert np.array_equal(res[i][0], counts) and np.array_equal(res[i][1], inf), i


def test_batch_beyond_the_slot_limit_is_refused():
    p = Params.from_row(ROW)
    b = PandemicBatch([p] * 15, list(range(1, 16)))
    with pytest.raises(Exception, match="exceeds the supported total"):
        b.setup()
    b.close()


def test_sweep_resimulation_batch_matches_single_runs():
    from pandemicthrust_amd import ensemble as E
    vecs = E.sample_scenarios(2, 303)
    params, ivs, seeds = [], [], []
    for i, v in enumerate(vecs):
        p, iv = E.scenario_inputs(v, 50_000_000, 7, seeds_per_million=True)
        for r in range(2):
            params.append(p)
            ivs.append(iv)
            seeds.append(5_000_000 + 2 * i + r)
    print("batch: 2 scenarios x 2 replicates x 50M agents, 7 days", flush=True)
    b = PandemicBatch(params, seeds)
    b.set_interventions(ivs)
    b.setup()
    res = b.run()
    b.close()
    for j in range(len(params)):
        print("single run %d" % j, flush=True)
        s = PandemicSim(params[j], seeds[j])
        s.set_interventions(ivs[j])
        s.setup_sim()
        counts, inf = s.run()
        s.close()
        assert np.array_equal(res[j][0], counts) and np.array_equal(res[j][1], inf), j
