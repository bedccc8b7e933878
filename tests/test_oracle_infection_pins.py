"""Pins of the oracle's infection arithmetic and winner rule (a18/a19:
PandemicSim.cu:3106-3233, :2307-2326, :3974-3988, :2828-2840, :3048-3103).

Each pin checks what the model's definition fixes (tests/infection_pins.py),
not a second copy of the oracle's code.  The last test builds on copies of the
oracle with one deliberate mistake each (OR_PLANT in oracle/pt_oracle.cpp:
a >= for the profile threshold, a shedding day off by one, the wrong xor
word, a wrong contact normalisation, a wrong strain factor, a generation not
incremented, slots in reverse order, no xor) and shows that every one of
them fails at least one pin.
"""
import numpy as np
import pytest

import oracle_ctypes as O
from infection_pins import InfectionPins, calibrated, reference_shed

# 471k agents, 15 days: two strains circulating together, so that enough
# slots see both strains succeed; pct and a away from 1 and 0 so that every
# term of the strain factor matters.
PIN_ROW = ["15", "2.5", "2.5", "500", "500", "0.2", "0.2", "0.7", "0.4"]
C0 = ["30", "1.8", "1.3", "10", "10", "0.0425", "0.0425", "1", "0"]


def traced_run(path, row, seed=1, days=None):
    """Steps the oracle day by day with the trace on; returns the pins."""
    p = O.params_from_row(row)
    sim = O.OracleSim(p, seed, path)
    sim.setup()
    sim.enable_trace(True)
    pins = InfectionPins(p)
    before = [sim.array(i) for i in range(6)]
    for d in range(days if days is not None else p.max_days):
        rc, _, _ = sim.step()
        after = [sim.array(i) for i in range(6)]
        if rc != 0:
            break
        inf, vic, typ, act = sim.trace()
        draw, prof = sim.trace_draws()
        pins.add_day(d, before, (inf, vic, typ, act, draw, prof), after)
        before = after
    return pins, sim


def profile_kat_failures(path):
    """device_getInfectionProfile at the threshold: y = (float)r / 2^32 and
    the profile is asymptomatic iff y > pct (strictly)."""
    out = []
    cases = [
        # (pct, r, expected profile)
        (0.5, 1 << 31, 2 + (1 << 31) % 3),               # y == 0.5 exactly: symptomatic
        (0.5, (1 << 31) - 1, 2 + ((1 << 31) - 1) % 3),   # rounds up to 2^31: y == 0.5
        (0.5, (1 << 31) + 128, 2 + ((1 << 31) + 128) % 3),  # tie rounds to even 2^31
        (0.5, (1 << 31) + 129, 5 + ((1 << 31) + 129) % 3),  # next float: y > 0.5
        (1.0, 0xFFFFFFFF, 2 + 0xFFFFFFFF % 3),           # rounds to 2^32: y == 1, symptomatic
        (0.0, 0, 2),                                     # y == 0: symptomatic
        (0.0, 1, 5 + 1),                                 # y == 2^-32 > 0
        (0.25, 1 << 30, 2 + (1 << 30) % 3),
        (0.25, (1 << 30) + 65, 5 + ((1 << 30) + 65) % 3),
    ]
    for pct, r, want in cases:
        row = list(C0)
        row[7] = repr(pct)
        sim = O.OracleSim(O.params_from_row(row), 1, path)
        got = sim.infection_profile(r)
        if got != want:
            out.append("profile(r=%#x, pct=%g) = %d, expected %d" % (r, pct, got, want))
    return out


def probability_kat_failures(path):
    """device_calculateInfectionProbability: shed[prof][dd] * factor / max(kval_sum, 1)
    over the reference's literals, with factor = R / ((1 - a) * pct)."""
    out = []
    row = ["30", "1.8", "1.3", "10", "10", "0.0425", "0.0425", "0.6", "0.25"]
    p = O.params_from_row(row)
    sim = O.OracleSim(p, 1, path)
    shed, fac = calibrated(p)
    for prof in range(6):
        for dd in range(10):
            for k in (0, 1, 3, 8):
                want = np.float32(shed[prof, dd] * fac[prof % 2]) / np.float32(max(k, 1))
                got = np.float32(sim.infection_probability(prof, dd, prof % 2, k))
                if got != want:
                    out.append("p(prof %d, day %d, strain %d, kval_sum %d) = %r, expected %r"
                               % (prof, dd, prof % 2, k, got, want))
                    return out
    return out


def special_case_failures(path):
    """Whole runs whose outcome the definition fixes."""
    out = []
    # pct = 1, a = 0 (the reference's default): y > 1 never holds, so every
    # infection is symptomatic: statuses <= 4, asymptomatic columns 0.
    sim = O.OracleSim(O.params_from_row(C0), 1, path)
    sim.setup()
    counts, _ = sim.run()
    if counts[:, 5:8].any() or counts[:, 13:16].any():
        out.append("pct = 1: asymptomatic statuses appear")
    if sim.array(O.OR_STATUS_P).max() > 4 or sim.array(O.OR_STATUS_S).max() > 4:
        out.append("pct = 1: a status above 4")
    # pct = 0, a = 0.5: the factor is R / 0 = inf, so every real contact of a
    # shedding infector succeeds, and every profile draw with r > 0 is
    # asymptomatic: the symptomatic columns hold only the seeds (profile 2)
    # until they recover on day 10.
    row = ["14", "1.8", "1.3", "10", "10", "0.0425", "0.0425", "0", "0.5"]
    sim = O.OracleSim(O.params_from_row(row), 1, path)
    sim.setup()
    counts, _ = sim.run()
    for d in range(len(counts)):
        want_p = [10 if d < 10 else 0, 0, 0]
        want_s = [10 if d < 10 else 0, 0, 0]
        if list(counts[d, 2:5]) != want_p or list(counts[d, 10:13]) != want_s:
            out.append("pct = 0, day %d: symptomatic columns %s %s" % (d, counts[d, 2:5], counts[d, 10:13]))
            break
    if counts[3, 5:8].sum() < 1000:
        out.append("pct = 0: the epidemic did not take off (%s)" % counts[3])
    return out


def r_calibration(path, r_value, seeds=300):
    """Seeds' successful draws over their 10 infectious days at pct = 1, a = 0:
    the strain factor is R and each symptomatic shedding row sums to 1, so a
    seed succeeds R * sum(shed[0]) times on average (less only on days with no
    contact at all).  Returns the pins and the finished simulation."""
    row = ["11", repr(r_value), repr(r_value), str(seeds), str(seeds), "1", "1", "1", "0"]
    pins, sim = traced_run(path, row)
    return pins, sim


def all_pin_failures(path):
    f = profile_kat_failures(path) + probability_kat_failures(path) + special_case_failures(path)
    pins, _ = traced_run(path, PIN_ROW)
    f += pins.failures()
    return f, pins


def test_infection_pins_hold_for_the_oracle():
    f, pins = all_pin_failures(O.LIB_PATH)
    assert not f, f
    # the run exercises what the pins need
    assert pins.days == 15 and len(pins.cell_z()) > 200
    assert pins.double[0] > 1000, pins.double
    assert pins.winners > 500000
    assert min(pins.prof_after_success.sum(1)) > 100000


@pytest.mark.parametrize("r_value", [0.5, 1.8, 3.0])
def test_r_calibration_closed_form(r_value):
    pins, sim = r_calibration(O.LIB_PATH, r_value)
    assert not pins.failures(), pins.failures()
    shed_sum = float(reference_shed()[0].astype(np.float64).sum())
    assert abs(shed_sum - 1.0) < 1e-5
    for s in range(2):
        o, e, v = pins.gen0[s]
        closed = 300 * r_value * shed_sum
        # the expected count is R * sum(shed) per seed up to seed-days without any contact
        assert 0.97 * closed <= e <= closed * (1 + 1e-6), (s, e, closed)
        assert abs(o - closed) <= 5 * np.sqrt(v) + 0.03 * closed, (s, o, closed)
    # output_rn.csv row 0 (final_countReproduction, PandemicSim.cu:2282-2295):
    # generation-1 agents per seed, all infected by the seeds in days 0..9
    cp, cs = sim.generation_offsets()
    for s, c in enumerate((cp, cs)):
        gen0, gen1 = c[1] - c[0], c[2] - c[1]
        assert gen0 == 300
        o = pins.gen0[s][0]
        assert 0 < gen1 <= o          # applied infections are a subset of the successful draws
        assert gen1 >= 0.7 * o        # household repeats waste some of them, not most


@pytest.mark.parametrize("k", range(1, 9))
def test_every_planted_mistake_fails_a_pin(k):
    path = O.plant_path(k)
    assert O.lib(path).or_plant() == k
    f, _ = all_pin_failures(path)
    assert f, "planted mistake %d passed every pin" % k
