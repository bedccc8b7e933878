"""Pins of the infection arithmetic (device_processContacts,
PandemicSim.cu:3106-3233; device_calculateInfectionProbability, :2307-2326;
device_getInfectionProfile, :3974-3988; setup_calculateInfectionData,
:2828-2840) and of the one-winner rule (device_setInfectionStatus,
:3048-3067; thrust_functors.h:96-122).

The checks do not restate the simulator's code path: they test what the
model's definition fixes about the simulation's observable behaviour, so a
plausible mistake in the restatement fails one of them.

- Per-contact rate: a slot with a real contact (kval 1) of an infector who is
  infectious for strain s on shedding day dd succeeds with probability
  shed[prof][dd] * R_s / ((1 - a) * pct) / max(kval_sum, 1), the shedding
  literals taken from the reference source (tests/golden/ref_tables.csv).
  Successes are counted per (strain, profile, day, kval_sum) cell and
  compared with the Poisson-binomial mean and variance.
- Slots with no contact and strains the infector does not shed never succeed.
- Profile draws: the symptomatic share is pct and the three profiles of
  each kind are equally likely; a success of one strain on a slot does not
  change that strain's profile share (the profile word is xored with the
  other strain's word); on a slot where both strains succeed with
  p <= 1/8 the pandemic profile is symptomatic, because the xor of two words
  below 2^29 is below 2^29 (y <= 1/8 < pct).
- Winner: for every victim and strain, the victim is infected today iff some
  slot's draw succeeded on it while it was susceptible; the winner is the
  successful slot with the smallest (position, slot); the victim takes that
  slot's profile draw, day d + 1 and the infector's generation + 1, and only
  that slot reports the infection.

Used by the oracle's CPU tests and by the GPU tests with the same inputs.
"""
import csv
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CT_NONE = 0


def reference_shed():
    """The 6x10 shedding literals of setup_loadParameters (float32)."""
    shed = np.zeros((6, 10), np.float32)
    with open(os.path.join(HERE, "golden", "ref_tables.csv")) as f:
        for r in csv.DictReader(f):
            if r["name"] == "shed":
                shed[int(r["i"]), int(r["j"])] = np.float32(float(r["value"]))
    return shed


def calibrated(params):
    """Shedding table and strain factors per setup_calculateInfectionData
    (PandemicSim.cu:2828-2840): asymptomatic rows 3..5 scaled by a, factor =
    R / ((1 - a) * pct), all in float32."""
    f32 = np.float32
    shed = reference_shed()
    a = f32(params.asymp_factor)
    shed[3:] = shed[3:] * a
    d = (f32(1.0) - a) * f32(params.percent_symptomatic)
    with np.errstate(divide="ignore", invalid="ignore"):
        fac = np.array([f32(params.r_pandemic) / d, f32(params.r_seasonal) / d], np.float32)
    return shed, fac


class InfectionPins:
    """Accumulates one simulation's contact days and checks them.

    add_day() takes the six state arrays before the day (status, day, gen per
    strain, after the previous day), the day's trace (infector, victim, type,
    action, draw bits, profile nibbles per slot, C slots per infector) and the
    state arrays after the day.
    """

    def __init__(self, params):
        self.params = params
        self.shed, self.fac = calibrated(params)
        self.pct = np.float32(params.percent_symptomatic)
        self.cells = {}            # (strain, prof, dd, kval_sum) -> [n, O, E, V]
        self.errors = []           # exact-rule violations (first few kept)
        self.n_errors = 0
        self.profiles = np.zeros((2, 8), np.int64)          # all slots: profile draws per strain
        self.prof_after_success = np.zeros((2, 2), np.int64)  # [strain] -> [symptomatic, asymptomatic]
        self.double = [0, 0, 0]    # double-success slots with p <= 1/8: n, pandemic asymptomatic, seasonal asymptomatic
        self.gen0 = np.zeros((2, 3))  # per strain: observed, expected, variance of gen-0 infectors' successes
        self.winners = 0
        self.days = 0

    def _fail(self, msg):
        self.n_errors += 1
        if len(self.errors) < 10:
            self.errors.append(msg)

    def add_day(self, d, before, trace, after):
        infector, victim, typ, action, draw, prof = trace
        n = len(victim)
        if n == 0:
            return
        C = 5 if d % 7 >= 5 else 8
        assert n % C == 0
        I = n // C
        self.days += 1
        typ = typ.reshape(I, C)
        real = typ != CT_NONE
        kval_sum = real.sum(1)
        inf_ids = infector.reshape(I, C)[:, 0]
        assert np.all(infector.reshape(I, C) == inf_ids[:, None])
        st = (before[0], before[1])
        dy = (before[2], before[3])
        gn = (before[4], before[5])
        draw = draw.reshape(I, C)
        prof = prof.reshape(I, C)
        action = action.reshape(I, C)
        vic = victim.reshape(I, C)
        infectious, pr, ddc, probs = [], [], [], []
        for s in range(2):
            stat = st[s][inf_ids].astype(np.int64)
            dd = d - dy[s][inf_ids].astype(np.int64)
            inf_s = (stat >= 2) & (dd >= 0) & (dd < 10)
            pr_s = np.where(inf_s, stat - 2, 0)
            dd_s = np.clip(dd, 0, 9)
            with np.errstate(invalid="ignore", over="ignore"):
                avg = self.shed[pr_s, dd_s] * self.fac[s]
                p = (avg / np.maximum(kval_sum, 1).astype(np.float32)).astype(np.float64)
            infectious.append(inf_s)
            pr.append(pr_s)
            ddc.append(dd_s)
            probs.append(np.where(inf_s, p, -1.0))
        for s in range(2):
            bit = 2 if s == 0 else 1
            ok = (draw & bit) != 0
            p = probs[s]
            # no contact, or not shedding: never a success
            bad = ok & ~(real & infectious[s][:, None])
            if bad.any():
                self._fail("day %d strain %d: %d successes on empty slots or non-infectious infectors"
                           % (d, s, int(bad.sum())))
            # certain outcomes: p*kval > 1 always succeeds (y <= 1 < p), p <= 0 or NaN never
            pk = np.where(real, p[:, None], 0.0)
            with np.errstate(invalid="ignore"):
                sure = real & (pk > 1.0)
                never = real & ~(pk > 0.0)
                sel = real & (pk > 0.0) & (pk < 1.0)
            if (ok & never).any() or (~ok & sure).any():
                self._fail("day %d strain %d: outcome contradicts p <= 0 or p > 1" % (d, s))
            # per-cell Poisson-binomial tallies over real contacts of infectious infectors
            rows, cols = np.nonzero(sel)
            if len(rows):
                keys = (pr[s][rows] * 10 + ddc[s][rows]) * 16 + kval_sum[rows]
                pv = pk[rows, cols]
                ov = ok[rows, cols].astype(np.float64)
                uk, inv = np.unique(keys, return_inverse=True)
                nn = np.bincount(inv)
                oo = np.bincount(inv, ov)
                ee = np.bincount(inv, pv)
                vv = np.bincount(inv, pv * (1 - pv))
                for j, k in enumerate(uk):
                    key = (s, int(k) // 160, (int(k) // 16) % 10, int(k) % 16)
                    c = self.cells.setdefault(key, [0, 0.0, 0.0, 0.0])
                    c[0] += int(nn[j]); c[1] += oo[j]; c[2] += ee[j]; c[3] += vv[j]
                # generation-0 infectors (the seeds): their successes over the infectious window
                g0 = gn[s][inf_ids[rows]] == 0
                self.gen0[s] += [ov[g0].sum(), pv[g0].sum(), (pv[g0] * (1 - pv[g0])).sum()]
            # profile draws of every slot
            pn = (prof >> (0 if s == 0 else 4)) & 0xF
            if ((pn < 2) | (pn > 7)).any():
                self._fail("day %d strain %d: profile draw outside 2..7" % (d, s))
            self.profiles[s] += np.bincount(pn.ravel(), minlength=16)[:8]
            # a success of this strain while the infector does not shed the
            # other strain: the word xored in (the other strain's, or the next
            # slot's pandemic word for the seasonal draw) is unconditioned, so
            # the asymptomatic share stays 1 - pct
            m = ok & ~infectious[1 - s][:, None]
            asym = (pn >= 5) & m
            self.prof_after_success[s] += [int(m.sum() - asym.sum()), int(asym.sum())]
        # both strains succeed on one slot with both p <= 1/8
        both = ((draw & 3) == 3) & (probs[0][:, None] <= 0.125) & (probs[1][:, None] <= 0.125)
        if both.any():
            pp = prof[both] & 0xF
            ps = prof[both] >> 4
            self.double[0] += int(both.sum())
            self.double[1] += int((pp >= 5).sum())
            self.double[2] += int((ps >= 5).sum())
        self._winners(d, before, after, inf_ids, vic, draw, prof, action, C)

    def _winners(self, d, before, after, inf_ids, vic, draw, prof, action, C):
        N = len(before[0])
        I = len(inf_ids)
        key = np.arange(I * C, dtype=np.int64).reshape(I, C)   # = position * C + slot: canonical order
        for s in range(2):
            bit = 2 if s == 0 else 1
            st0, st1 = before[s], after[s]
            dy1, gn0, gn1 = after[2 + s], before[4 + s], after[4 + s]
            v = vic.ravel()
            ok = ((draw.ravel() & bit) != 0) & (v >= 0)
            vv = np.where(ok, v, 0)
            ok &= st0[vv] == 1
            won = np.full(N, np.iinfo(np.int64).max, np.int64)
            np.minimum.at(won, v[ok], key.ravel()[ok])
            hit = won != np.iinfo(np.int64).max
            newly = (st0 == 1) & (st1 >= 2)
            if not np.array_equal(hit, newly):
                self._fail("day %d strain %d: %d victims infected without a winning draw or vice versa"
                           % (d, s, int((hit != newly).sum())))
                continue
            ids = np.nonzero(hit)[0]
            self.winners += len(ids)
            wk = won[ids]
            wpos = wk // C
            pn = ((prof.ravel()[wk] >> (0 if s == 0 else 4)) & 0xF).astype(np.uint8)
            if not np.array_equal(st1[ids], pn):
                self._fail("day %d strain %d: victim profile is not the winning slot's draw" % (d, s))
            if not np.all(dy1[ids] == (d + 1) & 0xFF):
                self._fail("day %d strain %d: infection day is not d + 1" % (d, s))
            want_gen = (gn0[inf_ids[wpos]].astype(np.int64) + 1) & 0xFF
            if not np.array_equal(gn1[ids].astype(np.int64), want_gen):
                self._fail("day %d strain %d: victim generation is not the winner's + 1" % (d, s))
            reported = (action.ravel() & bit) != 0
            want = np.zeros(I * C, bool)
            want[wk] = True
            if not np.array_equal(reported, want):
                self._fail("day %d strain %d: %d slots report an infection they did not win"
                           % (d, s, int((reported != want).sum())))

    # ---- results ----
    def cell_z(self, min_expected=10.0):
        out = []
        for key, (n, o, e, v) in sorted(self.cells.items()):
            if e >= min_expected and n - e >= min_expected and v > 0:
                out.append((key, n, o, e, (o - e) / math.sqrt(v)))
        return out

    def failures(self, z_max=6.0):
        """Every pin that fails, as text (empty when all hold)."""
        f = list(self.errors)
        if self.n_errors > len(self.errors):
            f.append("... %d exact-rule violations in all" % self.n_errors)
        cz = self.cell_z()
        if not cz:
            f.append("no calibration cell has enough contacts")
        else:
            worst = max(cz, key=lambda t: abs(t[4]))
            if abs(worst[4]) > z_max:
                f.append("per-contact rate off in cell (strain, profile, day, kval_sum) = %s: "
                         "%d contacts, %d successes, %.1f expected (z = %.1f)"
                         % (worst[0], worst[1], worst[2], worst[3], worst[4]))
            chi2 = sum(t[4] ** 2 for t in cz)
            k = len(cz)
            zchi = (chi2 - k) / math.sqrt(2 * k)
            if zchi > z_max:
                f.append("per-contact rates: chi2 %.0f over %d cells (z = %.1f)" % (chi2, k, zchi))
        for s in range(2):
            o, e, v = self.gen0[s]
            if v > 0 and abs(o - e) / math.sqrt(v) > z_max:
                f.append("strain %d: seeds' successes %d vs %.1f expected" % (s, o, e))
            n = self.profiles[s].sum()
            if n:
                asym = self.profiles[s][5:8].sum()
                q = 1.0 - float(self.pct)
                z = (asym - n * q) / math.sqrt(max(n * q * (1 - q), 1e-9))
                if abs(z) > z_max and 0 < q < 1:
                    f.append("strain %d: asymptomatic profile share %.5f, expected %.5f (z = %.1f)"
                             % (s, asym / n, q, z))
                if q <= 0 and asym:
                    f.append("strain %d: %d asymptomatic profile draws with pct = 1" % (s, asym))
                thirds = self.profiles[s][2:5] + self.profiles[s][5:8]
                chi = float(((thirds - n / 3.0) ** 2 / (n / 3.0)).sum())
                if chi > 40:   # chi2 with 2 degrees of freedom: p < 2e-9
                    f.append("strain %d: profiles within a kind not uniform (chi2 %.1f)" % (s, chi))
            sym, asy = self.prof_after_success[s]
            m = sym + asy
            q = 1.0 - float(self.pct)
            if m and 0 < q < 1:
                z = (asy - m * q) / math.sqrt(m * q * (1 - q))
                if abs(z) > z_max:
                    f.append("strain %d: asymptomatic share after a success %.4f of %d, expected %.4f (z = %.1f)"
                             % (s, asy / m, m, q, z))
        nd, pa, _ = self.double
        if float(self.pct) > 0.125 and pa:
            f.append("%d of %d double-success slots drew an asymptomatic pandemic profile" % (pa, nd))
        return f
