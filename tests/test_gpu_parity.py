"""HIP path vs the CPU oracle, bit-exact (integer, byte and index results).

Every test calls the gfx950 library through its C ABI (pandemicthrust_amd.api)
and the oracle through oracle/build/liboracle.so on the same inputs.
"""
import numpy as np
import pytest

import oracle_ctypes as O

pytestmark = pytest.mark.gpu

C0 = ["30", "1.8", "1.3", "10", "10", "0.0425", "0.0425", "1", "0"]


def _sims(row, seed):
    from pandemicthrust_amd import Params, PandemicSim
    gpu = PandemicSim(Params.from_row(row), seed)
    ref = O.OracleSim(O.params_from_row(row), seed)
    n = gpu.setup_sim()
    ref.setup()
    assert n == ref.n_people
    return gpu, ref


def _assert_state(gpu, ref, tag=""):
    names = ["status_p", "status_s", "day_p", "day_s", "gen_p", "gen_s"]
    for which, name in enumerate(names):
        g, r = gpu.array(which), ref.array(which)
        if not np.array_equal(g, r):
            bad = np.nonzero(g != r)[0]
            raise AssertionError("%s %s differs at %d agents, first %s: gpu %s oracle %s" % (
                tag, name, len(bad), bad[:5], g[bad[:5]], r[bad[:5]]))


def test_setup_tables_bit_exact():
    from pandemicthrust_amd.api import PT_AGE, PT_HH_OFFSETS, PT_HH_TYPES, PT_HOUSEHOLD, PT_WP_OFFSETS, PT_WP_PEOPLE
    gpu, ref = _sims(C0, 1)
    for g, r in [(PT_HH_TYPES, O.OR_HH_TYPES), (PT_HH_OFFSETS, O.OR_HH_OFFSETS), (PT_HOUSEHOLD, O.OR_HOUSEHOLD),
                 (PT_AGE, O.OR_AGE), (PT_WP_OFFSETS, O.OR_WP_OFFSETS), (PT_WP_PEOPLE, O.OR_WP_PEOPLE)]:
        a, b = gpu.array(g), ref.array(r)
        assert a.shape == b.shape and np.array_equal(a, b), (g, np.nonzero(a != b)[0][:5])
    sc = gpu.scalars()
    assert sc[0] == ref.scalar(0) and sc[1] == ref.scalar(1) and sc[2] == ref.scalar(2)
    assert sc[3] == ref.scalar(3) and sc[4] == ref.scalar(4)
    _assert_state(gpu, ref, "after seeding")


def test_day_by_day_bit_exact():
    """Each day: counts, infected list, the day's full errand table and all state."""
    from pandemicthrust_amd.api import PT_ERRAND_OFFSETS, PT_ERRAND_PEOPLE, PT_INFECTED
    gpu, ref = _sims(C0, 1)
    gpu.set_table_mode(1)
    for day in range(14):  # two weeks: every weekday and weekend layout
        rg, cg, ig = gpu.step()
        rr, cr, ir = ref.step()
        assert (rg, ig) == (rr, ir), (day, rg, ig, rr, ir)
        assert np.array_equal(cg, cr), (day, cg, cr)
        if rr == 1:
            break
        assert np.array_equal(gpu.array(PT_INFECTED), ref.array(O.OR_INFECTED)), day
        assert np.array_equal(gpu.array(PT_ERRAND_OFFSETS), ref.array(O.OR_ERRAND_OFFSETS)), day
        assert np.array_equal(gpu.array(PT_ERRAND_PEOPLE), ref.array(O.OR_ERRAND_PEOPLE)), day
        assert gpu.scalars()[3] == ref.scalar(3), day
        _assert_state(gpu, ref, "day %d" % day)


@pytest.mark.parametrize("row,seed", [
    (C0, 1),
    (C0, 7),
    (["60", "2.4", "1.6", "25", "5", "0.0425", "0.0425", "0.6", "0.4"], 3),   # asymptomatic profiles
    (["40", "1.8", "1.3", "10", "10", "0.3", "0.3", "1", "0"], 1),            # ~700k agents
    (["20", "3.0", "2.0", "3", "3", "0.0005", "0.0005", "1", "0"], 11),       # 500 households, min-1 locations
    (["45", "1.8", "1.3", "10", "10", "0.0425", "0.5", "1", "0"], 2),          # many small locations
    (["20", "1.8", "1.3", "0", "0", "0.0425", "0.0425", "1", "0"], 4),          # nobody infected: ends on day 0
    (["30", "1.8", "1.3", "0", "12", "0.0425", "0.0425", "1", "0"], 5),         # seasonal strain only
    (["30", "2.0", "1.3", "15", "0", "0.0425", "0.0425", "1", "0"], 6),         # pandemic strain only
])
def test_full_run_bit_exact(row, seed):
    gpu, ref = _sims(row, seed)
    cg, ig = gpu.run()
    cr, ir = ref.run()
    assert np.array_equal(ig, ir), (ig, ir)
    assert np.array_equal(cg, cr)
    _assert_state(gpu, ref, "end")
    gp, gs = gpu.generation_offsets()
    rp, rs = ref.generation_offsets()
    assert np.array_equal(gp, rp) and np.array_equal(gs, rs)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_table_modes_give_identical_days(mode):
    """Full tables, queried-bucket tables and the automatic choice produce
    the same state every day (the tables differ only in unqueried buckets)."""
    row = ["45", "2.2", "1.5", "20", "20", "0.1", "0.1", "0.8", "0.5"]
    gpu, ref = _sims(row, 9)
    gpu.set_table_mode(mode)
    for day in range(45):
        rg, cg, ig = gpu.step()
        rr, cr, ir = ref.step()
        assert (rg, ig) == (rr, ir) and np.array_equal(cg, cr), day
        _assert_state(gpu, ref, "mode %d day %d" % (mode, day))
        if rr == 1:
            break


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_one_pass_tables_give_identical_days(mode):
    """4,707 agents in 31 locations: the weekday keys (< 64 buckets) are
    sorted in one radix pass, whose digit is the whole key (the offsets then
    come from that pass's histogram alone), and the weekend keys in two; each
    table path, every day against the oracle."""
    row = ["20", "1.8", "1.3", "5", "5", "0.002", "0.002", "1", "0"]
    gpu, ref = _sims(row, 3)
    gpu.set_table_mode(mode)
    for day in range(20):
        rg, cg, ig = gpu.step()
        rr, cr, ir = ref.step()
        assert (rg, ig) == (rr, ir) and np.array_equal(cg, cr), day
        _assert_state(gpu, ref, "mode %d day %d" % (mode, day))
        if rr == 1:
            break


def test_zero_reproduction_closed_form():
    """R = 0: no draw succeeds; the seeds recover on day 10 and the run
    stops on day 10 with everybody susceptible or recovered."""
    row = ["30", "0", "0", "10", "10", "0.0425", "0.0425", "1", "0"]
    gpu, _ = _sims(row, 1)
    counts, inf = gpu.run()
    n = gpu.n_people
    assert len(inf) == 11 and list(inf[:10]) == [20] * 10 and inf[10] == 0
    for d in range(10):
        assert counts[d, 1] == n - 10 and counts[d, 2] == 10 and counts[d, 0] == 0
    assert counts[10, 0] == 10 and counts[10, 1] == n - 10


def test_rerun_after_restore_is_identical():
    gpu, _ = _sims(C0, 5)
    gpu.snapshot()
    c1, i1 = gpu.run()
    s1 = [gpu.array(w).copy() for w in range(6)]
    gpu.restore()
    c2, i2 = gpu.run()
    assert np.array_equal(c1, c2) and np.array_equal(i1, i2)
    for w in range(6):
        assert np.array_equal(s1[w], gpu.array(w))
