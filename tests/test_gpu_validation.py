"""Validation mode on the GPU (the reference's SIM_VALIDATION checks,
debug.cu:138-442, :847-950).

The per-contact trace (infector, victim, type, action) must equal the
oracle's trace day by day; the traced contact locations must satisfy the
reference's own contact checks (home: both in the infector's household;
workplace and school: both at the infector's workplace; after-school: both
at the infector's after-school site; errands: both go to that location in
that hour); the workplace and errand hooks must equal the oracle's
re-generation (host_recalcWorkplace / host_recalc_* semantics).  The host
program's validation files are checked for format and for an error-free
debug.txt.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_ctypes as O
from pandemicthrust_amd import PandemicSim, Params
from pandemicthrust_amd.api import PT_AGE, PT_ERRANDS, PT_HOUSEHOLD, PT_LOCATION_TYPE_OFFSETS, PT_WORKPLACE

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "pandemicthrust_amd", "bin", "pandemicthrust")
HOME, WORK, SCHOOL, ERRAND, AFTER, NONE = 5, 1, 2, 3, 4, 0


def _check_locations(t, N, W, hh, wp, er, weekend):
    inf, vic, typ, loc = t["infector"], t["victim"], t["type"], t["loc"]
    assert ((vic == -1) == (loc == -1)).all()            # unused slots carry no location
    m = (typ == NONE) & (vic >= 0)                         # nobody else at the location
    assert (vic[m] <= 1).all()
    m = typ == HOME
    assert (loc[m] == hh[inf[m]]).all() and (hh[vic[m]] == hh[inf[m]]).all()
    m = (typ == WORK) | (typ == SCHOOL)
    assert (loc[m] == wp[inf[m]]).all() and (wp[vic[m]] == wp[inf[m]]).all()
    if weekend:
        assert not ((typ == WORK) | (typ == SCHOOL) | (typ == AFTER)).any()
        e3 = er.reshape(N, 3)
        m = np.nonzero(typ == ERRAND)[0]
        assert all(loc[i] in e3[inf[i]] and loc[i] in e3[vic[i]] for i in m)
    else:
        h0, h1 = er[:N], er[N:]
        m = typ == AFTER
        assert (loc[m] == h0[inf[m]]).all() and (h0[vic[m]] == loc[m]).all()
        m = np.nonzero(typ == ERRAND)[0]
        for i in m:
            hour = loc[i] // W
            row = h0 if hour == 0 else h1
            assert row[inf[i]] == loc[i] and row[vic[i]] == loc[i], i


@pytest.mark.parametrize("row,seed,mode", [
    (["21", "1.8", "1.3", "10", "10", "0.0425", "0.0425", "1", "0"], 1, 0),
    (["21", "1.8", "1.3", "10", "10", "0.0425", "0.0425", "1", "0"], 1, 1),   # full tables every day
    (["21", "1.8", "1.3", "10", "10", "0.0425", "0.0425", "1", "0"], 1, 2),   # queried buckets only, every day
    (["16", "2.4", "1.6", "30", "20", "0.03", "0.06", "0.8", "0.4"], 5, 0),
])
def test_trace_matches_oracle_with_consistent_locations(row, seed, mode):
    gpu = PandemicSim(Params.from_row(row), seed)
    gpu.set_table_mode(mode)
    gpu.enable_trace()
    N = gpu.setup_sim()
    ref = O.OracleSim(O.params_from_row(row), seed)
    ref.enable_trace(True)
    ref.setup()
    W = int(gpu.scalars()[2])
    hh, age, wp = gpu.array(PT_HOUSEHOLD), gpu.array(PT_AGE), gpu.array(PT_WORKPLACE)
    assert np.array_equal(wp, ref.array(O.OR_WP_KEYS))
    # location-type offsets: a scan of the type counts ending at W; children attend a school of their age's
    # type (CHILD_AGE_SCHOOLTYPE_LOOKUP_HOST, types 3..7)
    offs = gpu.array(PT_LOCATION_TYPE_OFFSETS)
    assert len(offs) == 15 and offs[0] == 0 and offs[-1] == W and (np.diff(offs) >= 0).all()
    child = age < 5
    t = age[child].astype(np.int64) + 3
    assert ((wp[child] >= offs[t]) & (wp[child] < offs[t + 1])).all()
    days = 0
    for day in range(int(row[0])):
        rc, c, i = gpu.step()
        rrc, rcnt, ri = ref.step()
        assert i == ri and np.array_equal(c, rcnt), day
        if i == 0 or rc == 1:
            break
        t = gpu.trace()
        oi, ov, ot, oa = ref.trace()
        assert np.array_equal(t["infector"], oi), day
        assert np.array_equal(t["victim"], ov), day
        assert np.array_equal(t["type"], ot), day
        assert np.array_equal(t["action"], oa), day
        weekend = day % 7 >= 5
        er = gpu.array(PT_ERRANDS)
        assert len(er) == (3 if weekend else 2) * N
        # the errand hook against the oracle's re-generation, a sample of agents
        sc = gpu.scalars()
        ref.set_offsets(int(sc[4]), int(sc[5]))
        for idx in np.linspace(0, N - 1, 97).astype(int):
            if weekend:
                assert tuple(er[3 * idx:3 * idx + 3]) == ref.recalc_weekend(int(idx)), (day, idx)
            else:
                d1, d2, _ = ref.recalc_weekday(int(idx), int(age[idx]))   # d2 carries hour 1: W + location
                assert er[idx] == d1 and er[N + idx] == d2, (day, idx)
        _check_locations(t, N, W, hh, wp, er, weekend)
        days += 1
    assert days >= 10
    gpu.close()


def test_tracing_leaves_results_unchanged():
    row = ["30", "1.8", "1.3", "10", "10", "0.0425", "0.0425", "1", "0"]
    a = PandemicSim(Params.from_row(row), 2)
    a.setup_sim()
    ca, ia = a.run()
    b = PandemicSim(Params.from_row(row), 2)
    b.enable_trace()
    b.setup_sim()
    cb, ib = b.run()
    assert np.array_equal(ca, cb) and np.array_equal(ia, ib)
    for k in range(6):
        assert np.array_equal(a.array(k), b.array(k))
    a.close()
    b.close()


def test_host_program_validation_files(tmp_path):
    d = str(tmp_path)
    with open(os.path.join(d, "constants.csv"), "w") as f:
        f.write("constants,max_days,reproduction_number_pandemic,reproduction_number_seasonal,initial_infected_pandemic,"
                "initial_infected_seasonal,people_scaling_factor,location_scaling_factor,percent_symptomatic,"
                "asymp_factor\nconstants,12,1.8,1.3,10,10,0.0425,0.0425,1,0\n")
    with open(os.path.join(d, "seed.txt"), "w") as f:
        f.write("1\n")
    r = subprocess.run([EXE, "--validation"], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    inf = open(os.path.join(d, "debug_infected.csv")).read()
    assert inf == "current_day, i, idx, status_p, day_p, gen_p, status_s, day_s, gen_s\n"
    con = open(os.path.join(d, "debug_contacts.csv")).read().splitlines()
    assert con[0] == ("current_day, i, infector_idx, victim_idx, contact_type, contact_loc, infector_loc, "
                      "victim_loc, locs_matched")
    pat = re.compile(r"^\d+,\d+,\d+,-?\d+,CONTACT_TYPE_[A-Z]+,-?\d+,-?\d+,-?\d+,[01]$")
    assert len(con) > 100 and all(pat.match(l) for l in con[1:]), [l for l in con[1:] if not pat.match(l)][:3]
    assert all(l.endswith(",1") for l in con[1:])
    act = open(os.path.join(d, "debug_actions.csv")).read().splitlines()
    assert act[0] == "current_day, i, infector, victim, action_type, action_type_string"
    assert len(act) == len(con)
    pat = re.compile(r"^\d+,\d+,\d+,-?\d+,([0-3]),(NONE|SEAS|PAND|BOTH)$")
    names = {"0": "NONE", "1": "SEAS", "2": "PAND", "3": "BOTH"}
    for l in act[1:]:
        m = pat.match(l)
        assert m and names[m[1]] == m[2], l
    dbg = open(os.path.join(d, "debug.txt")).read()
    assert "ERROR" not in dbg and "day 0\ninfected: 20" in dbg
