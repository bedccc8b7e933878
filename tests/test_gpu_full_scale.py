"""The 50M-agent workload (constants,100,1.8,1.3,10,10,21.24,21.24,1,0, seed 1)
at its full size.

- Setup: every population table equals the oracle's, bit-exact (the oracle's
  setup takes a few seconds at this size).
- The first four days equal the oracle's day by day (counts, infected list,
  all agent state), and so does the first week under every intervention at
  once (latent and infectious periods, a school closure over days 2-5,
  workplace and errand reductions, vaccination; 3,000 seeds).
- The day's errand tables, built by the radix sort over 100M (weekday) and
  150M (weekend) pairs, checked through size-independent properties against
  the errand hook (people_errands): every entry sits in the bucket of its
  person's destination for that hour, every person appears exactly 2 (3)
  times, ids ascend inside every bucket, offsets run from 0 to the pair
  count.
- The full 100-day run: status counts conserve N every day, susceptibles
  never grow and recovered never shrink, the infected count lies between
  the larger strain count and their sum, and at the end every agent's
  status, day and generation agree (susceptible <=> day and generation
  unset; infected and recovered carry a generation no later than their day).
"""
import numpy as np
import pytest

import oracle_ctypes as O
from pandemicthrust_amd import PandemicSim, Params
from pandemicthrust_amd.api import (PT_AGE, PT_ERRAND_OFFSETS, PT_ERRAND_PEOPLE, PT_ERRANDS, PT_HH_OFFSETS,
                                    PT_HH_TYPES, PT_HOUSEHOLD, PT_INFECTED, PT_WP_OFFSETS, PT_WP_PEOPLE)

pytestmark = pytest.mark.gpu

ROW = ["100", "1.8", "1.3", "10", "10", "21.24", "21.24", "1", "0"]
SEED = 1


def _state(sim):
    return [sim.array(k) for k in range(6)]


def test_setup_and_first_days_equal_the_oracle():
    gpu = PandemicSim(Params.from_row(ROW), SEED)
    n = gpu.setup_sim()
    ref = O.OracleSim(O.params_from_row(ROW), SEED)
    ref.setup()
    assert n == ref.n_people > 49_000_000
    for g, r in [(PT_HH_TYPES, O.OR_HH_TYPES), (PT_HH_OFFSETS, O.OR_HH_OFFSETS), (PT_HOUSEHOLD, O.OR_HOUSEHOLD),
                 (PT_AGE, O.OR_AGE), (PT_WP_OFFSETS, O.OR_WP_OFFSETS), (PT_WP_PEOPLE, O.OR_WP_PEOPLE)]:
        a, b = gpu.array(g), ref.array(r)
        assert a.shape == b.shape and np.array_equal(a, b), g
    for day in range(4):
        rg, cg, ig = gpu.step()
        rr, cr, ir = ref.step()
        assert (rg, ig) == (rr, ir) and np.array_equal(cg, cr), day
        assert np.array_equal(gpu.array(PT_INFECTED), ref.array(O.OR_INFECTED)), day
        for k, (a, b) in enumerate(zip(_state(gpu), [ref.array(w) for w in range(6)])):
            assert np.array_equal(a, b), (day, k)
    gpu.close()


def test_first_week_with_interventions_equals_the_oracle():
    from pandemicthrust_amd import Interventions
    row = ["100", "1.8", "1.3", "2000", "1000", "21.24", "21.24", "1", "0"]
    iv = Interventions(latent=(1, 0), infectious=(8, 6), school_close_start=2, school_close_days=4,
                       workplace_reduction=0.3, errand_reduction=0.2, vaccination=(0.15, 0.05))
    gpu = PandemicSim(Params.from_row(row), SEED)
    gpu.set_interventions(iv)
    ref = O.OracleSim(O.params_from_row(row), SEED)
    ref.set_interventions(iv)
    n = gpu.setup_sim()
    ref.setup()
    assert n == ref.n_people > 49_000_000
    for day in range(7):   # five weekdays (closure on days 2-5) and a weekend
        print("interventions at 50M: day %d" % day, flush=True)
        rg, cg, ig = gpu.step()
        rr, cr, ir = ref.step()
        assert (rg, ig) == (rr, ir) and np.array_equal(cg, cr), day
        assert np.array_equal(gpu.array(PT_INFECTED), ref.array(O.OR_INFECTED)), day
        for k, (a, b) in enumerate(zip(_state(gpu), [ref.array(w) for w in range(6)])):
            assert np.array_equal(a, b), (day, k)
    assert ig > 3000      # the epidemic grew under the interventions
    gpu.close()


def _check_table(gpu, N, W, hours, per):
    off = gpu.array(PT_ERRAND_OFFSETS).astype(np.int64)
    people = gpu.array(PT_ERRAND_PEOPLE).astype(np.int64)
    er = gpu.array(PT_ERRANDS).astype(np.int64)
    assert len(off) == hours * W + 1 and off[0] == 0 and off[-1] == len(people) == per * N
    cnt = np.diff(off)
    assert (cnt >= 0).all()
    bucket = np.repeat(np.arange(hours * W, dtype=np.int64), cnt)
    # every person exactly `per` times
    assert (np.bincount(people, minlength=N) == per).all()
    # ascending ids inside every bucket
    same = bucket[1:] == bucket[:-1]
    assert (np.diff(people)[same] > 0).all()
    # each entry in its person's destination bucket for that hour
    if per == 2:
        hour = bucket // W
        dest = np.where(hour == 0, er[people], er[N + people])
        assert np.array_equal(dest, bucket)
    else:
        e3 = er.reshape(N, 3)[people]
        assert ((e3 == bucket[:, None]).sum(axis=1) >= 1).all()
    return len(people)


def test_full_size_errand_tables():
    gpu = PandemicSim(Params.from_row(ROW), SEED)
    N = gpu.setup_sim()
    W = int(gpu.scalars()[2])
    gpu.set_table_mode(1)  # full tables every day
    rc, c, i = gpu.step()  # day 0, a weekday
    assert rc == 0 and _check_table(gpu, N, W, 2, 2) == 2 * N
    for day in range(1, 6):
        rc, c, i = gpu.step()
        assert rc == 0, day
    assert _check_table(gpu, N, W, 10, 3) == 3 * N  # day 5, a weekend day
    gpu.close()


def test_full_run_invariants():
    gpu = PandemicSim(Params.from_row(ROW), SEED)
    N = gpu.setup_sim()
    counts, inf = gpu.run()
    assert len(inf) == 100
    c = counts.astype(np.int64)
    assert (c[:, 0:8].sum(axis=1) == N).all() and (c[:, 8:16].sum(axis=1) == N).all()
    assert (np.diff(c[:, 1]) <= 0).all() and (np.diff(c[:, 9]) <= 0).all()     # susceptible
    assert (np.diff(c[:, 0]) >= 0).all() and (np.diff(c[:, 8]) >= 0).all()     # recovered
    ip, is_ = c[:, 2:8].sum(axis=1), c[:, 10:16].sum(axis=1)
    assert (inf >= np.maximum(ip, is_)).all() and (inf <= ip + is_).all()
    st_p, st_s, dy_p, dy_s, gn_p, gn_s = _state(gpu)
    for st, dy, gn in ((st_p, dy_p, gn_p), (st_s, dy_s, gn_s)):
        sus = st == 1
        assert ((dy == 255) == sus).all() and ((gn == 255) == sus).all()
        assert (gn[~sus] <= dy[~sus]).all() and (dy[~sus] <= 100).all()
        assert (st <= 7).all()
    # the last counts row precedes the last day's infections
    assert (st_p == 1).sum() <= c[-1, 1] and (st_s == 1).sum() <= c[-1, 9]
    gpu.close()
