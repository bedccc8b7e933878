"""The infection pins (tests/infection_pins.py) on the HIP path, through the
C ABI, with the oracle stepped beside it: every traced slot's draw outcome,
profile draw and reported action equal the oracle's, and the pins that a
plausible mistake fails hold on the GPU's own results (a18/a19:
PandemicSim.cu:3106-3233, :2307-2326, :3974-3988, :2828-2840, :3048-3103).
"""
import numpy as np
import pytest

import oracle_ctypes as O
from infection_pins import InfectionPins, reference_shed
from test_oracle_infection_pins import C0, PIN_ROW

pytestmark = pytest.mark.gpu


def traced_pair(row, seed=1):
    """GPU and oracle stepped together with the trace on; returns the GPU's
    pins after checking every day's trace against the oracle's."""
    from pandemicthrust_amd import Params, PandemicSim
    p = Params.from_row(row)
    gpu = PandemicSim(p, seed)
    ref = O.OracleSim(O.params_from_row(row), seed)
    gpu.enable_trace(True)
    ref.enable_trace(True)
    n = gpu.setup_sim()
    ref.setup()
    assert n == ref.n_people
    pins = InfectionPins(p)
    before = [gpu.array(i) for i in range(6)]
    for d in range(p.max_days):
        rg, cg, ig = gpu.step()
        rr, cr, ir = ref.step()
        assert (rg, ig) == (rr, ir) and np.array_equal(cg, cr), (d, rg, ig, rr, ir)
        after = [gpu.array(i) for i in range(6)]
        for i in range(6):
            assert np.array_equal(after[i], ref.array(i)), (d, i)
        if rg != 0:
            break
        t = gpu.trace()
        draw, prof = gpu.trace_draws()
        inf_r, vic_r, typ_r, act_r = ref.trace()
        draw_r, prof_r = ref.trace_draws()
        assert np.array_equal(t["infector"], inf_r) and np.array_equal(t["victim"], vic_r), d
        assert np.array_equal(t["type"], typ_r) and np.array_equal(t["action"], act_r), d
        assert np.array_equal(draw, draw_r), (d, np.nonzero(draw != draw_r)[0][:5])
        assert np.array_equal(prof, prof_r), (d, np.nonzero(prof != prof_r)[0][:5])
        pins.add_day(d, before, (t["infector"], t["victim"], t["type"], t["action"], draw, prof), after)
        before = after
    return pins, gpu


def test_gpu_infection_pins_and_trace_equal_the_oracle():
    print("case: pins row %s seed 1" % ",".join(PIN_ROW), flush=True)
    pins, _ = traced_pair(PIN_ROW)
    f = pins.failures()
    assert not f, f
    assert pins.days == 15 and pins.double[0] > 1000 and pins.winners > 500000


def test_gpu_r_calibration_closed_form():
    row = ["11", "1.8", "1.8", "300", "300", "1", "1", "1", "0"]
    print("case: R calibration row %s" % ",".join(row), flush=True)
    pins, gpu = traced_pair(row)
    assert not pins.failures(), pins.failures()
    closed = 300 * 1.8 * float(reference_shed()[0].astype(np.float64).sum())
    cp, cs = gpu.generation_offsets()
    for s, c in enumerate((cp, cs)):
        o, e, v = pins.gen0[s]
        assert 0.97 * closed <= e <= closed * (1 + 1e-6)
        assert abs(o - closed) <= 5 * np.sqrt(v) + 0.03 * closed
        assert c[1] - c[0] == 300 and 0.7 * o <= c[2] - c[1] <= o


@pytest.mark.parametrize("pct,asymp", [(1.0, 0.0), (0.0, 0.5)])
def test_gpu_profile_special_cases(pct, asymp):
    """pct = 1: no asymptomatic status ever; pct = 0 (factor R / 0 = inf):
    every real contact of a shedding infector succeeds and only the seeds are
    symptomatic.  Whole runs, GPU equal to the oracle."""
    from pandemicthrust_amd import Params, PandemicSim
    row = list(C0)
    row[0] = "14" if pct == 0.0 else "30"
    row[7], row[8] = repr(pct), repr(asymp)
    print("case: special row %s" % ",".join(row), flush=True)
    gpu = PandemicSim(Params.from_row(row), 1)
    gpu.setup_sim()
    cg, ig = gpu.run()
    ref = O.OracleSim(O.params_from_row(row), 1)
    ref.setup()
    cr, ir = ref.run()
    assert np.array_equal(cg, cr) and np.array_equal(ig, ir)
    if pct == 1.0:
        assert not cg[:, 5:8].any() and not cg[:, 13:16].any()
        assert gpu.array(0).max() <= 4 and gpu.array(1).max() <= 4
    else:
        for d in range(len(cg)):
            seeds = 10 if d < 10 else 0
            assert list(cg[d, 2:5]) == [seeds, 0, 0] and list(cg[d, 10:13]) == [seeds, 0, 0], d
        assert cg[3, 5:8].sum() > 1000
