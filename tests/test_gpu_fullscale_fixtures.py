"""Whole runs of the large workloads against committed oracle fixtures
(tests/golden/make_fullscale_fixture.py writes them on the CPU), and the
256-scenario replicate set batched against its solo runs.

- 50M agents x 100 days (constants,100,1.8,1.3,10,10,21.24,21.24,1,0, seed 1):
  the host program's output_infected_stats.csv, output_rn.csv and
  infected_max.txt byte-identical to the oracle's; the C-ABI run's daily rows,
  generation offsets and the SHA-256 of all six final per-agent arrays equal
  the oracle's (runToCompletion, PandemicSim.cu:877-967; the files of
  :2771-2818, :2282-2295, :863).
- The same run saved at day 60 (pt_sim_checkpoint_save) and resumed in a
  fresh handle: identical to the fixture.
- The largest rows of the replicate set (about 2M agents x 180 days): the
  same, solo and as replicates of the full 256-row batch.
- All 256 rows batched: every replicate equal to its own solo run.
"""
import glob
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

from pandemicthrust_amd import PandemicBatch, PandemicSim, Params
from pandemicthrust_amd import outputs
from pandemicthrust_amd.scenarios import load_replicate_set

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "pandemicthrust_amd", "bin", "pandemicthrust")
FILES = ["output_infected_stats.csv", "output_rn.csv", "infected_max.txt"]
HEADER = ("constants,max_days,reproduction_number_pandemic,reproduction_number_seasonal,initial_infected_pandemic,"
          "initial_infected_seasonal,people_scaling_factor,location_scaling_factor,percent_symptomatic,asymp_factor")
NAMES = ["status_p", "status_s", "day_p", "day_s", "gen_p", "gen_s"]


def fixture(name):
    d = os.path.join(GOLDEN, name)
    with open(os.path.join(d, "state.json")) as f:
        st = json.load(f)
    files = {}
    for n in FILES:
        with open(os.path.join(d, n)) as f:
            files[n] = f.read()
    return st, files


def _row(st):
    return st["constants"].split(",")[1:]


def _check_run(tag, st, files, counts, inf, n_people, cp, cs, arrays):
    row = _row(st)
    max_days = int(row[0])
    assert n_people == st["n_people"], (tag, n_people)
    assert len(inf) == st["days"], (tag, len(inf), st["days"])
    assert list(cp) == st["gen_offsets_p"] and list(cs) == st["gen_offsets_s"], tag
    for name, a in zip(NAMES, arrays):
        assert hashlib.sha256(a.tobytes()).hexdigest() == st["sha256"][name], (tag, name)
    assert outputs.infected_stats_text(counts, inf) == files["output_infected_stats.csv"], tag
    assert outputs.infected_max_text(inf, n_people) == files["infected_max.txt"], tag
    assert outputs.rn_text(cp, cs, max_days) == files["output_rn.csv"], tag


def _solo_check(name):
    st, files = fixture(name)
    print("case: %s %s seed %d" % (name, st["constants"], st["seed"]), flush=True)
    sim = PandemicSim(Params.from_row(_row(st)), st["seed"])
    n = sim.setup_sim()
    counts, inf = sim.run()
    cp, cs = sim.generation_offsets()
    _check_run(name, st, files, counts, inf, n, cp, cs, [sim.array(k) for k in range(6)])
    sim.close()


def test_50M_agents_100_days_equal_the_oracle():
    _solo_check("loop_50M_100d")


def test_50M_agents_resumed_mid_run_equal_the_oracle():
    """Saved at the day-60 boundary (rising towards the peak), loaded into a
    fresh handle and run to the end: the same files, tallies and final
    state as the oracle's uninterrupted run."""
    st, files = fixture("loop_50M_100d")
    print("case: loop_50M_100d resumed from day 60", flush=True)
    a = PandemicSim(Params.from_row(_row(st)), st["seed"])
    n = a.setup_sim()
    for _ in range(60):
        a.step()
    blob = a.save_checkpoint()
    a.close()
    b = PandemicSim(Params.from_row(_row(st)), st["seed"])
    b.setup_only()
    b.load_checkpoint(blob)
    del blob
    b.run()
    counts, inf = b.daily_rows()
    cp, cs = b.generation_offsets()
    _check_run("loop_50M_100d resumed", st, files, counts, inf, n, cp, cs, [b.array(k) for k in range(6)])
    b.close()


def test_50M_agents_host_program_files_equal_the_oracle(tmp_path):
    st, files = fixture("loop_50M_100d")
    d = str(tmp_path)
    with open(os.path.join(d, "constants.csv"), "w") as f:
        f.write(HEADER + "\n" + st["constants"] + "\n")
    with open(os.path.join(d, "seed.txt"), "w") as f:
        f.write("%d\n" % st["seed"])
    r = subprocess.run([EXE], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    for n in FILES:
        with open(os.path.join(d, n)) as f:
            assert f.read() == files[n], n


def _replicate_fixtures():
    return sorted(os.path.basename(p) for p in glob.glob(os.path.join(GOLDEN, "replicate_row*")))


@pytest.mark.parametrize("name", _replicate_fixtures())
def test_largest_scenarios_equal_the_oracle(name):
    _solo_check(name)


def test_256_scenario_batch_equals_solo_runs_and_the_oracle():
    rows = load_replicate_set()
    assert len(rows) == 256
    print("case: 256-row replicate batch", flush=True)
    b = PandemicBatch([Params.from_row(f) for _, f in rows], [s for s, _ in rows])
    n = b.setup()
    res = b.run()
    fixtures = {int(name[len("replicate_row"):]): fixture(name) for name in _replicate_fixtures()}
    assert len(fixtures) >= 2
    for r, (seed, f) in enumerate(rows):
        c, inf = res[r]
        cp, cs = b.generation_offsets(r)
        arrays = [b.array(r, k) for k in range(6)]
        if r in fixtures:
            st, files = fixtures[r]
            assert _row(st) == list(f) and st["seed"] == seed
            _check_run("batch replicate %d" % r, st, files, c, inf, int(n[r]), cp, cs, arrays)
        solo = PandemicSim(Params.from_row(f), seed)
        assert solo.setup_sim() == n[r]
        sc, sinf = solo.run()
        assert np.array_equal(inf, sinf) and np.array_equal(c, sc), r
        sp, ss = solo.generation_offsets()
        assert np.array_equal(cp, sp) and np.array_equal(cs, ss), r
        for k in range(6):
            assert np.array_equal(arrays[k], solo.array(k)), (r, NAMES[k])
        solo.close()
    b.close()
