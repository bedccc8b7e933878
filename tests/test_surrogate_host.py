"""Host logic of the surrogate and its training set, without a GPU: the
strain-swap augmentation, curve statistics, screening flags and threshold
calibration (on a stub ensemble whose predictions are set by hand), the
nearest-training-row distance and the convex-hull test, and the training
set's part files."""
import math

import numpy as np
import pytest

from pandemicthrust_amd import ensemble as E
from pandemicthrust_amd import surrogate as S


def test_augment_appends_strain_swapped_twins():
    rng = np.random.default_rng(1)
    X = E.sample_scenarios(12, 5)
    Y = rng.random((12, 2, 120, 3))
    Xa, Ya = S.augment(X, Y)
    assert Xa.shape == (24, 16) and Ya.shape == (24, 2, 120, 3)
    assert np.array_equal(Xa[:12], X) and np.array_equal(Ya[:12], Y)
    for a, b in E.STRAIN_PAIRS:
        assert np.array_equal(Xa[12:, a], X[:, b]) and np.array_equal(Xa[12:, b], X[:, a])
    shared = [j for j in range(16) if all(j not in p for p in E.STRAIN_PAIRS)]
    assert shared == [8, 9, 10, 11, 14, 15]
    assert np.array_equal(Xa[12:, shared], X[:, shared])
    assert np.array_equal(Ya[12:, 0], Y[:, 1]) and np.array_equal(Ya[12:, 1], Y[:, 0])
    assert np.array_equal(E.swap_strains(E.swap_strains(X)), X)


def test_curve_stats_and_errors():
    inc = np.zeros((2, 120))
    inc[0, 30] = 99.0
    inc[0, 10:20] = 9.0
    inc[1, 50] = 9.0
    d, h, a = S.curve_stats(np.log1p(inc))
    assert list(d) == [30, 50] and np.allclose(h, [99, 9]) and np.allclose(a, [99 + 90, 9])
    e = S.curve_errors(np.log1p(inc * 1.1), np.log1p(inc))
    assert np.allclose(e["peak_height"], 0.1) and np.allclose(e["attack"], 0.1)
    assert np.all(e["peak_day"] == 0)
    # scaled prediction: the predicted curve times `scale` is compared
    e2 = S.curve_errors(np.log1p(inc / 50), np.log1p(inc), pred_scale=50.0)
    assert np.allclose(e2["peak_height"], 0, atol=1e-9) and np.allclose(e2["attack"], 0, atol=1e-9)


class Stub:
    """predict(Xn) -> members x n x 720: member k predicts base + offset[k]
    on the log-mean channel of strain 0."""

    def __init__(self, base, offsets):
        self.base, self.offsets = base, offsets

    def predict(self, Xn):
        n = len(Xn)
        out = np.zeros((len(self.offsets), n, 2, 120, 3), np.float32)
        for k, off in enumerate(self.offsets):
            out[k, :, 0, :, 0] = self.base[:n, None] + off[:n, None]
        return out.reshape(len(self.offsets), n, 720)


def test_nearest_distance_is_brute_force():
    rng = np.random.default_rng(2)
    T, Q = rng.random((300, 16)), rng.random((50, 16))
    brute = np.sqrt(((Q[:, None, :] - T[None, :, :]) ** 2).sum(-1)).min(1)
    assert np.allclose(S.nearest_distance(Q, T, chunk=7), brute)


def test_in_hull():
    T = np.vstack([np.zeros(16), np.eye(16)])     # simplex corners
    assert S.in_hull(np.full(16, 0.5 / 16), T)
    assert not S.in_hull(np.full(16, 0.2), T)        # coordinates sum to 3.2 > 1
    assert not S.in_hull(-0.01 * np.ones(16), T)


def test_calibrate_takes_quantiles_of_held_out_spread_and_distance():
    rng = np.random.default_rng(3)
    train = E.sample_scenarios(200, 11)
    held = E.sample_scenarios(100, 12)
    offs = [rng.random(100) * s for s in (0.0, 1.0, 2.0)]
    stub = Stub(np.ones(100), offs)
    th = S.calibrate(stub, held, train, quantile=0.9)
    spread = S.member_spread(stub.predict(S.normalize(held)))
    assert math.isclose(th["disagreement"], float(np.quantile(spread, 0.9)), rel_tol=1e-12)
    dist = S.nearest_distance(S.normalize(held), S.normalize(train))
    assert math.isclose(th["distance"], float(np.quantile(dist, 0.9)), rel_tol=1e-12)


def test_screen_flags():
    train = E.sample_scenarios(400, 21)
    # inside the training box, except the row planted below
    X = np.clip(E.sample_scenarios(300, 22), train.min(axis=0), train.max(axis=0))
    n = len(X)
    base = np.linspace(0.0, 3.0, n)                  # the last rows have the largest attack
    offs = [np.zeros(n), np.zeros(n)]
    offs[1][5] = 4.0                                 # members disagree on row 5
    X[7, 0] = 0.5                                    # outside the training box in p_pandemic
    stub = Stub(base, offs)
    th = {"disagreement": 0.5, "distance": 10.0}
    sc = S.screen(stub, X, train, th, top=0.01)
    assert sc["disagree"][5] and sc["disagree"].sum() == 1
    assert sc["outside"][7] and sc["outside"].sum() == 1
    k = math.ceil(0.01 * n)
    assert sc["top"].sum() == k and set(np.nonzero(sc["top"])[0]) == set(range(n - k, n))
    assert np.array_equal(sc["resimulate"], sc["disagree"] | sc["outside"] | sc["top"])
    # a distance threshold below every row's distance sends every row back
    sc2 = S.screen(stub, X, train, {"disagreement": 0.5, "distance": 0.0}, top=0.01)
    assert sc2["outside"].all()


def test_training_set_parts_merge_in_order(tmp_path):
    X = E.sample_scenarios(10, 31)
    Y = np.random.default_rng(4).random((10, 2, 120, 3)) * 10
    info = {"seconds": 1.5, "person_days": 1000}
    p0, p1, out = tmp_path / "a.npz", tmp_path / "b.npz", tmp_path / "m.npz"
    E.save_part(p1, X[6:], Y[6:], 6, info, 64, 1000000, 120, 7)
    E.save_part(p0, X[:6], Y[:6], 0, info, 64, 1000000, 120, 7)
    E.merge_parts([p1, p0], out)
    Xm, Ym, meta = E.load_training_set(out)
    assert np.array_equal(Xm, X)
    assert np.array_equal(Ym, Y.astype(np.float16).astype(np.float32))
    assert meta["replicates"] == 64 and meta["seconds"] == 3.0 and meta["person_days"] == 2000
    with pytest.raises(AssertionError):
        E.merge_parts([p1], tmp_path / "bad.npz")


def test_committed_training_set_is_the_stated_one():
    X, Y, meta = E.load_training_set(S.DATA)
    assert X.shape == (8192, 16) and Y.shape == (8192, 2, 120, 3)
    assert meta["replicates"] == 64 and meta["agents"] == 1_000_000 and meta["days"] == 120
    assert np.array_equal(X, E.sample_scenarios(8192, E.TRAINING_SET_SEED))
    assert np.isfinite(Y).all() and Y.min() >= 0.0
