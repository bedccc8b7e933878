"""Pins the CPU oracle against vectors produced by the reference's own code.

The fixtures in tests/golden/ come from tests/golden/make_golden.py, which
compiles the reference's host_functions.cpp and vendored Random123
threefry.h (from /root/reference) into a harness.  Each check here would fail
on a plausible restatement mistake: a wrong rotation or key-injection
constant (Threefry), a swapped union word (recalc slots), a 32-bit counter
(the 'wrap' config crosses 2^32), a '<' for '>' in a walk or a missed
subtraction (fish functions at CDF boundaries), a wrong scaling rounding.
"""
import csv
import os

import numpy as np
import pytest

import oracle_ctypes as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rows(name):
    with open(os.path.join(GOLD, name)) as f:
        return list(csv.DictReader(f))


# Published Random123 known-answer vectors (kat_vectors of the Random123
# distribution, 20 rounds), verified against the vendored header by the
# first rows of threefry_ref.csv.
KAT_2x64 = [
    ((0, 0, 0, 0), (0xC2B6E3A8C2C69865, 0x6F81ED42F350084D)),
    ((2**64 - 1,) * 4, (0xE02CB7C4D95D277A, 0xD06633D0893B8B68)),
    ((0x243F6A8885A308D3, 0x13198A2E03707344, 0xA4093822299F31D0, 0x082EFA98EC4E6C89),
     (0x263C7D30BB0F0AF1, 0x56BE8361D3311526)),
]
KAT_2x32 = [
    ((0, 0, 0, 0), (0x6B200159, 0x99BA4EFE)),
    ((2**32 - 1,) * 4, (0x1CB996FC, 0xBB002BE7)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xC4923A9C, 0x483DF7A0)),
]


def test_threefry_published_kat():
    for (c0, c1, k0, k1), out in KAT_2x64:
        assert O.threefry2x64(c0, c1, k0, k1) == out
    for (c0, c1, k0, k1), out in KAT_2x32:
        assert O.threefry2x32(c0, c1, k0, k1) == out


def test_threefry_matches_vendored_header():
    n = 0
    for r in rows("threefry_ref.csv"):
        args = [int(r[k]) for k in ("c0", "c1", "k0", "k1")]
        exp = (int(r["o0"]), int(r["o1"]))
        got = O.threefry2x64(*args) if r["kind"] == "2x64" else O.threefry2x32(*args)
        assert got == exp, r
        n += 1
    assert n > 150


def test_glibc_seed_sequence():
    # srand(1) on glibc: the first draws are fixed by glibc's TYPE_3 generator.
    assert O.glibc_seeds(1)[:2] == [1804289383, 846930886]


def test_model_tables_match_reference_source():
    tabs = {}
    for r in rows("ref_tables.csv"):
        tabs.setdefault(r["name"], {})[(int(r["i"]), int(r["j"]))] = float(r["value"])
    shapes = {"wd_assign": (4, 2), "we_assign": (6, 2), "shed": (6, 10)}
    for name, entries in tabs.items():
        if name == "child_school":
            continue  # unused by the reference (PandemicSim.cu:2412-2413 uses row + 3)
        exp = np.zeros(shapes.get(name, (max(i for i, _ in entries) + 1, 1)))
        for (i, j), v in entries.items():
            exp[i, j] = v
        got = O.raw_table(name)
        np.testing.assert_array_equal(got.astype(np.float32), exp.reshape(-1).astype(np.float32), err_msg=name)


CFG = {}
for _r in rows("ref_scaling.csv"):
    CFG.setdefault(_r["cfg"], (float(_r["lscale"]), int(_r["W"])))
SEEDS = {r["cfg"]: int(r["core_seed"]) for r in rows("host_workplace.csv")}


def make_sim(cfg):
    lscale, _ = CFG[cfg]
    p = O.Params(30, 1.8, 1.3, 10, 10, 0.0425, lscale, 1.0, 0.0)
    return O.OracleSim(p, SEEDS[cfg])


@pytest.mark.parametrize("cfg", sorted(CFG))
def test_scaling_matches_reference(cfg):
    sim = make_sim(cfg)
    cnt, off = sim.type_tables()
    for r in rows("ref_scaling.csv"):
        if r["cfg"] != cfg:
            continue
        t = int(r["type"])
        assert cnt[t] == int(r["count"]) and off[t] == int(r["offset"]), r
    assert int(cnt.sum()) == CFG[cfg][1]


def test_scaling_closed_form_sizes():
    # W at the two benchmark configs (every type count rounded half up,
    # min 1 where the base count is > 0): 545 and 271,872.
    assert CFG["small_100k"][1] == 545
    assert CFG["full_50M"][1] == 271872


@pytest.mark.parametrize("cfg", sorted(CFG))
def test_recalc_workplace_golden(cfg):
    sim = make_sim(cfg)
    for r in rows("host_workplace.csv"):
        if r["cfg"] != cfg:
            continue
        sim.set_offsets(int(r["workplace_off"]), 0)
        assert sim.recalc_workplace(int(r["idx"]), int(r["age"])) == int(r["workplace"]), r


@pytest.mark.parametrize("cfg", sorted(CFG))
def test_recalc_weekday_golden(cfg):
    sim = make_sim(cfg)
    for r in rows("host_weekday.csv"):
        if r["cfg"] != cfg:
            continue
        sim.set_offsets(0, int(r["errand_off"]))
        d1, d2, prof = sim.recalc_weekday(int(r["idx"]), int(r["age"]))
        assert (d1, d2, prof) == (int(r["dest1"]), int(r["dest2"]), int(r["profile"])), r


@pytest.mark.parametrize("cfg", sorted(CFG))
def test_recalc_weekend_golden(cfg):
    sim = make_sim(cfg)
    for r in rows("host_weekend.csv"):
        if r["cfg"] != cfg:
            continue
        sim.set_offsets(0, int(r["errand_off"]))
        assert sim.recalc_weekend(int(r["idx"])) == (int(r["d0"]), int(r["d1"]), int(r["d2"])), r


@pytest.mark.parametrize("cfg", sorted(CFG))
def test_fish_functions_golden(cfg):
    sim = make_sim(cfg)
    n = 0
    for r in rows("host_fish.csv"):
        if r["cfg"] != cfg:
            continue
        u = int(r["r"])
        assert sim.fish_workplace(u) == int(r["workplace"]), r
        assert sim.fish_school_and_age(u) == (int(r["child_age"]), int(r["school"])), r
        assert sim.fish_weekday_dest(u, 5) == int(r["wd_adult"]), r
        assert sim.fish_weekday_dest(u, 0) == int(r["wd_child"]), r
        assert sim.fish_weekend_dest(u) == int(r["weekend"]), r
        n += 1
    assert n >= 700


def test_household_type_walk_is_the_inverse_cdf():
    """device_setup_fishHouseholdType (PandemicSim.cu:2360-2371) has no host
    twin in the reference, so the oracle's walk is pinned by what the
    mathematics fixes: for the reference's float32 CDF literals
    (ref_tables.csv, hh_cdf) and y = float(r) / float(0xFFFFFFFF), the walk
    returns the first row whose CDF value is >= y, clipped to the last row
    (numpy's searchsorted, side='left').  The draws include every r whose y
    equals a CDF value exactly (a '>=' for '>' fails there), its float32
    neighbours on both sides (an off-by-one row fails), the extremes and
    random draws."""
    cdf = np.array([float(r["value"]) for r in rows("ref_tables.csv") if r["name"] == "hh_cdf"], np.float32)
    assert len(cdf) == 9 and cdf[-1] == np.float32(1.0)
    scale = np.float32(np.uint32(0xFFFFFFFF))  # rounds to 2**32
    assert float(scale) == 2.0 ** 32
    rs = [0, 1, 2**31, 2**32 - 1, 2**32 - 128, 2**32 - 129]
    for c in cdf[:-1]:
        r0 = int(np.float64(c) * 2.0**32)  # exact: c * 2**32 is an integer (float32 c < 1)
        assert np.float32(r0) / scale == c
        rs += [r0 + d for d in range(-300, 301, 1)]
    rs += [int(x) for x in np.random.default_rng(7).integers(0, 2**32, 20000, dtype=np.uint64)]
    rs = np.array(sorted(set(r for r in rs if 0 <= r < 2**32)), np.uint64)
    y = rs.astype(np.float32) / scale
    expect = np.minimum(np.searchsorted(cdf, y, side="left"), len(cdf) - 1)
    got = np.array([O.fish_household_type(int(r)) for r in rs])
    bad = np.nonzero(got != expect)[0]
    assert len(bad) == 0, [(int(rs[i]), float(y[i]), int(got[i]), int(expect[i])) for i in bad[:5]]
    # every row is reachable and the exact ties land on the row they close
    assert set(got.tolist()) == set(range(len(cdf)))
    for i, c in enumerate(cdf[:-1]):
        assert O.fish_household_type(int(np.float64(c) * 2.0**32)) == i


def test_oracle_continues_from_a_saved_day():
    """or_set_state: a run continued from a copy of another run's state at a
    day boundary ends exactly where the uninterrupted run ends (the CPU
    baseline times peak days this way)."""
    import oracle_ctypes as O
    row = ["30", "1.8", "1.3", "10", "10", "0.0425", "0.0425", "1", "0"]
    full = O.OracleSim(O.params_from_row(row), 3)
    full.setup()
    c_all, i_all = full.run()
    a = O.OracleSim(O.params_from_row(row), 3)
    a.setup()
    for _ in range(12):
        a.step()
    state = [a.array(k) for k in range(6)]
    b = O.OracleSim(O.params_from_row(row), 3)
    b.setup()
    b.set_state(a.scalar(6), a.scalar(3), state)
    c_b, i_b = b.run()
    assert np.array_equal(c_b, c_all[12:]) and np.array_equal(i_b, i_all[12:])
    for k in range(6):
        assert np.array_equal(b.array(k), full.array(k)), k
