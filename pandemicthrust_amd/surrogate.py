"""Epidemic surrogate: an ensemble of MLPs trained on scenario ensembles
(include/pandemic/pt_surrogate.h), and surrogate screening of scenario
sweeps.

Inputs are scenario vectors (ensemble.SCENARIO_DIMS) scaled to [0, 1] per
dimension; outputs are ensemble.summarise() targets flattened to 720 values
(2 strains x 120 days x (mean, q10, q90) of log(1 + incidence)).
"""
import argparse
import ctypes as C
import json
import math
import os

import numpy as np

from .api import _check, lib
from .ensemble import SCENARIO_DIMS, swap_strains

IN, HIDDEN, OUT = 16, 512, 720
DIMS = [IN, HIDDEN, HIDDEN, HIDDEN, HIDDEN, OUT]
_LO = np.array([d[1] for d in SCENARIO_DIMS], np.float64)
_HI = np.array([d[2] for d in SCENARIO_DIMS], np.float64)


def normalize(X):
    """Scenario vectors -> [0, 1] per dimension (float32)."""
    return ((np.asarray(X, np.float64) - _LO) / (_HI - _LO)).astype(np.float32)


def targets(Y):
    """ensemble.run_ensemble() Y (n x 2 x 120 x 3) -> n x 720 float32."""
    Y = np.asarray(Y)
    assert Y.shape[1:] == (2, 120, 3), Y.shape
    return Y.reshape(len(Y), OUT).astype(np.float32)


def augment(X, Y):
    """Training rows with their strain-swapped twins appended: the scenario
    with its per-strain dimensions swapped (ensemble.swap_strains) has the
    strains' outcomes swapped, so its targets are Y with the strain axis
    reversed.  X: n x 16 raw vectors, Y: n x 2 x days x 3."""
    X, Y = np.asarray(X), np.asarray(Y)
    return np.concatenate([X, swap_strains(X)]), np.concatenate([Y, Y[:, ::-1]])


def glorot(seed, members):
    """Glorot-uniform weights and zero biases, one parameter block per member."""
    count = int(lib().pt_mlp_param_count())
    out = np.zeros((members, count), np.float32)
    for k in range(members):
        rng = np.random.default_rng([seed, k])
        off = 0
        for l in range(len(DIMS) - 1):
            a = math.sqrt(6.0 / (DIMS[l] + DIMS[l + 1]))
            n = DIMS[l] * DIMS[l + 1]
            out[k, off:off + n] = rng.uniform(-a, a, n)
            off += n + DIMS[l + 1]
    return out


def training_flops(batch, members=1):
    """Matrix-product flops of one training step (forward, input and weight
    gradients; the first layer needs no input gradient)."""
    macs = sum(DIMS[l] * DIMS[l + 1] for l in range(len(DIMS) - 1))
    macs_dx = sum(DIMS[l] * DIMS[l + 1] for l in range(1, len(DIMS) - 1))
    return 2.0 * batch * members * (2 * macs + macs_dx)


class Surrogate:
    """MLP ensemble on one GPU."""

    def __init__(self, members=5, batch=1024, seed=1, device=0):
        h = C.c_void_p()
        _check(lib().pt_mlp_create(members, batch, device, C.byref(h)))
        self.h, self.members, self.batch = h, members, batch
        self.count = int(lib().pt_mlp_param_count())
        self.set_params(glorot(seed, members))
        self.n_data = 0

    def close(self):
        if getattr(self, "h", None):
            lib().pt_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32).reshape(self.members, self.count)
        _check(lib().pt_mlp_set_params(self.h, p.ctypes.data))

    def params(self):
        p = np.zeros((self.members, self.count), np.float32)
        _check(lib().pt_mlp_get_params(self.h, p.ctypes.data))
        return p

    def set_data(self, X, Y):
        """X: n x 16 scaled inputs, Y: n x 720 targets."""
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        _check(lib().pt_mlp_set_data(self.h, X.ctypes.data, Y.ctypes.data, len(X)))
        self.n_data = len(X)

    def train(self, steps, lr=1e-3, seed=0):
        """Adam steps over epoch permutations of the training set; returns the
        batch losses (steps x members)."""
        rng = np.random.default_rng(seed)
        need = steps * self.batch
        parts, have = [], 0
        while have < need:
            parts.append(rng.permutation(self.n_data))
            have += self.n_data
        order = np.ascontiguousarray(np.concatenate(parts)[:need], np.int32)
        loss = np.zeros((steps, self.members), np.float32)
        _check(lib().pt_mlp_train(self.h, order.ctypes.data, steps, C.c_float(lr), loss.ctypes.data))
        return loss

    def use_graphs(self, on=True):
        """Training steps replayed from a HIP graph (default) or enqueued
        kernel by kernel; results are identical."""
        _check(lib().pt_mlp_use_graphs(self.h, 1 if on else 0))

    def last_train_ms(self):
        return float(lib().pt_mlp_last_train_ms(self.h))

    def gradients(self, rows):
        rows = np.ascontiguousarray(rows, np.int32)
        assert len(rows) == self.batch
        g = np.zeros((self.members, self.count), np.float32)
        loss = np.zeros(self.members, np.float32)
        _check(lib().pt_mlp_gradients(self.h, rows.ctypes.data, g.ctypes.data, loss.ctypes.data))
        return g, loss

    def predict(self, X):
        """members x n x 720."""
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros((self.members, len(X), OUT), np.float32)
        _check(lib().pt_mlp_predict(self.h, X.ctypes.data, len(X), out.ctypes.data))
        return out


def curve_stats(logmean, scale=1.0):
    """Peak day, peak height and attack (total incidence) of incidence
    curves given as log(1 + incidence) means (... x days), with incidence
    multiplied by `scale` (a population ratio)."""
    inc = np.expm1(np.clip(np.asarray(logmean, np.float64), 0.0, None)) * scale
    return inc.argmax(axis=-1), inc.max(axis=-1), inc.sum(axis=-1)


def curve_errors(pred_logmean, sim_logmean, pred_scale=1.0):
    """Relative errors of the surrogate's curves against the simulator's
    ensemble means (scenarios x days, log(1 + incidence)): peak day, peak
    height and attack, each |pred - sim| / sim (sim at least 1)."""
    pd, ph, pa = curve_stats(pred_logmean, pred_scale)
    sd, sh, sa = curve_stats(sim_logmean)
    return {"peak_day": np.abs(pd - sd) / np.maximum(sd, 1), "peak_height": np.abs(ph - sh) / np.maximum(sh, 1.0),
            "attack": np.abs(pa - sa) / np.maximum(sa, 1.0), "peak_day_abs": np.abs(pd - sd).astype(np.float64)}


def nearest_distance(Xn, train_n, chunk=4096):
    """Euclidean distance from each scaled row of Xn to its nearest training
    row (both in [0, 1]^16)."""
    T = np.asarray(train_n, np.float64)
    t2 = (T * T).sum(axis=1)
    out = np.empty(len(Xn))
    for i in range(0, len(Xn), chunk):
        x = np.asarray(Xn[i:i + chunk], np.float64)
        d2 = (x * x).sum(axis=1)[:, None] + t2[None, :] - 2.0 * x @ T.T
        out[i:i + chunk] = np.sqrt(np.maximum(d2.min(axis=1), 0.0))
    return out


def in_hull(x, train_n):
    """Whether the scaled point x lies in the convex hull of the scaled
    training rows: a feasible lambda >= 0 with sum 1 and train^T lambda = x
    (a linear program, scipy's HiGHS)."""
    from scipy.optimize import linprog
    T = np.asarray(train_n, np.float64)
    A = np.vstack([T.T, np.ones(len(T))])
    b = np.concatenate([np.asarray(x, np.float64), [1.0]])
    r = linprog(np.zeros(len(T)), A_eq=A, b_eq=b, bounds=(0, None), method="highs")
    return r.status == 0


def member_spread(P, days=120):
    """Mean over days and strains of the members' standard deviation of the
    predicted log-incidence mean (P: members x n x 720)."""
    return P.std(axis=0).reshape(P.shape[1], 2, days, 3)[..., 0].mean(axis=(1, 2))


def calibrate(surrogate, heldout_X, train_X, quantile=0.99):
    """Re-simulation thresholds from held-out in-distribution scenarios (raw
    vectors): the `quantile` of the members' spread and of the distance to the
    nearest training scenario."""
    Xn, Tn = normalize(heldout_X), normalize(train_X)
    spread = member_spread(surrogate.predict(Xn))
    return {"disagreement": float(np.quantile(spread, quantile)),
            "distance": float(np.quantile(nearest_distance(Xn, Tn), quantile)), "quantile": quantile}


def screen(surrogate, X, train_X, thresholds, top=0.01, days=120):
    """Surrogate screening of raw scenario vectors X (n x 16; train_X: the
    raw training vectors, both scaled here).  Returns the ensemble mean
    prediction, its spread, and which rows go back to the simulator:
    members disagree (spread above thresholds["disagreement"]), the row is
    out of distribution (outside the training ranges in some dimension, or
    farther from its nearest training scenario than
    thresholds["distance"]), or it is in the top `top` share by predicted
    pandemic attack (sum of predicted incidence over the first `days` days).

    Out of distribution is not tested against the convex hull of the
    training set: in 16 dimensions almost no new point lies inside the hull
    of a few thousand uniform samples (in_hull), so that test would send
    every row back; the nearest-training-scenario distance, calibrated on
    held-out scenarios, measures the same thing usefully."""
    Xn, Tn = normalize(X), normalize(train_X)
    P = surrogate.predict(Xn)
    mean = P.mean(axis=0)
    spread = member_spread(P)
    lo, hi = Tn.min(axis=0), Tn.max(axis=0)
    dist = nearest_distance(Xn, Tn)
    outside = ((Xn < lo) | (Xn > hi)).any(axis=1) | (dist > thresholds["distance"])
    disagree = spread > thresholds["disagreement"]
    attack = np.expm1(np.clip(mean.reshape(len(X), 2, 120, 3)[:, 0, :days, 0], 0, None)).sum(axis=1)
    k = max(1, int(math.ceil(top * len(X))))
    top_rows = np.zeros(len(X), bool)
    top_rows[np.argsort(-attack)[:k]] = True
    resim = disagree | outside | top_rows
    return {"mean": mean, "spread": spread, "distance": dist, "outside": outside, "disagree": disagree,
            "top": top_rows, "resimulate": resim, "attack": attack}


DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "surrogate_ensemble.npz")
SCHEDULE = ((8000, 1e-3), (8000, 3e-4), (8000, 1e-4))


def main():
    """Trains the 5-member ensemble on a training set (the last `heldout`
    scenarios held out, every other one also strain-swapped) and prints one
    JSON line: steps, ms per step, TFLOP/s, held-out median errors.

        python -m pandemicthrust_amd.surrogate [--data surrogate_ensemble.npz] [--scale 1.0] [--eager]

    --scale multiplies every phase of the learning-rate schedule's step
    counts (8000 steps each at 1e-3, 3e-4, 1e-4); --eager enqueues each step
    kernel by kernel instead of replaying a HIP graph (same results; for
    profilers that trace launches one by one)."""
    from .ensemble import load_training_set
    ap = argparse.ArgumentParser()
    ap.add_argument("--data", default=DATA)
    ap.add_argument("--heldout", type=int, default=256)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--eager", action="store_true")
    a = ap.parse_args()
    X, Y, meta = load_training_set(a.data)
    ntr = len(X) - a.heldout
    sur = Surrogate(members=5, batch=1024, seed=1)
    sur.use_graphs(not a.eager)
    Xa, Ya = augment(X[:ntr], Y[:ntr])
    sur.set_data(normalize(Xa), targets(Ya))
    steps, ms = 0, 0.0
    for n, lr in SCHEDULE:
        n = max(1, int(round(n * a.scale)))
        sur.train(n, lr=lr, seed=n)
        steps += n
        ms += sur.last_train_ms()
    P = sur.predict(normalize(X[ntr:])).mean(axis=0).reshape(len(X) - ntr, 2, Y.shape[2], 3)
    sur.close()
    out = {"steps": steps, "ms_per_step": round(ms / steps, 4),
           "tflops": round(training_flops(1024, 5) * steps / (ms * 1e-3) / 1e12, 2)}
    for k, name in enumerate(("pandemic", "seasonal")):
        e = curve_errors(P[:, k, :, 0], Y[ntr:, k, :, 0])
        out[name] = {m: round(float(np.median(v)), 4) for m, v in e.items()}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
