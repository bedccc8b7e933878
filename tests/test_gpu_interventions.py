"""Interventions on the GPU, bit-exact against the CPU oracle (single
simulations) and against single runs (replicate batch)."""
import numpy as np
import pytest

import oracle_ctypes as O

pytestmark = pytest.mark.gpu

STATE = ["status_p", "status_s", "day_p", "day_s", "gen_p", "gen_s"]
C0 = ["40", "1.8", "1.3", "10", "10", "0.0425", "0.0425", "1", "0"]


def _iv(**kw):
    from pandemicthrust_amd import Interventions
    return Interventions(**kw)


SETS = {
    "latent": dict(latent=(2, 4), infectious=(7, 10)),
    "closure": dict(school_close_start=3, school_close_days=9),
    "reductions": dict(workplace_reduction=0.4, errand_reduction=0.25),
    "vaccination": dict(vaccination=(0.3, 0.15)),
    "all": dict(latent=(1, 0), infectious=(8, 6), school_close_start=0, school_close_days=20, workplace_reduction=0.3,
                errand_reduction=0.5, vaccination=(0.2, 0.35)),
}


def _pair(row, seed, iv):
    from pandemicthrust_amd import Params, PandemicSim
    gpu = PandemicSim(Params.from_row(row), seed)
    gpu.set_interventions(iv)
    ref = O.OracleSim(O.params_from_row(row), seed)
    ref.set_interventions(iv)
    n = gpu.setup_sim()
    ref.setup()
    assert n == ref.n_people
    return gpu, ref


def _same_state(gpu, ref, tag):
    for which, name in enumerate(STATE):
        g, r = gpu.array(which), ref.array(which)
        if not np.array_equal(g, r):
            bad = np.nonzero(g != r)[0]
            raise AssertionError("%s: %s differs at %d agents, first %s" % (tag, name, len(bad), bad[:5]))


@pytest.mark.parametrize("name", sorted(SETS))
def test_interventions_match_oracle_day_by_day(name):
    gpu, ref = _pair(C0, 2, _iv(**SETS[name]))
    _same_state(gpu, ref, name + " after seeding")
    for day in range(40):
        rg, cg, ig = gpu.step()
        rr, cr, ir = ref.step()
        assert rg == rr and ig == ir and np.array_equal(cg, cr), (name, day)
        _same_state(gpu, ref, "%s day %d" % (name, day))
        if rg:
            break


@pytest.mark.parametrize("mode", [1, 2])
def test_interventions_all_table_modes(mode):
    gpu, ref = _pair(C0, 4, _iv(**SETS["all"]))
    gpu.set_table_mode(mode)
    cg, ig = gpu.run()
    cr, ir = ref.run()
    assert np.array_equal(ig, ir) and np.array_equal(cg, cr)
    _same_state(gpu, ref, "mode %d" % mode)


def test_batch_with_interventions_matches_single_runs():
    from pandemicthrust_amd import PandemicBatch, PandemicSim, Params
    rows = [(C0, 1), (["35", "2.4", "1.1", "5", "30", "0.021", "0.021", "1", "0"], 2),
            (["50", "2.0", "1.6", "7", "9", "0.03", "0.05", "0.9", "0.3"], 3)]
    ivs = [_iv(**SETS["all"]), _iv(), _iv(**SETS["latent"], vaccination=(0.1, 0.0))]
    b = PandemicBatch([Params.from_row(r) for r, _ in rows], [s for _, s in rows])
    b.set_interventions(ivs)
    b.setup()
    res = b.run()
    for r, ((row, seed), iv) in enumerate(zip(rows, ivs)):
        s = PandemicSim(Params.from_row(row), seed)
        s.set_interventions(iv)
        s.setup_sim()
        c, inf = s.run()
        assert np.array_equal(res[r][1], inf) and np.array_equal(res[r][0], c), r
        for which, name in enumerate(STATE):
            assert np.array_equal(b.array(r, which), s.array(which)), (r, name)
        s.close()
    b.close()
