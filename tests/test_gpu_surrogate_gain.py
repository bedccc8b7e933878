"""What the trained surrogate gains over the method without it.

On the last 256 scenarios of the committed training set
(pandemicthrust_amd/data/surrogate_ensemble.npz: 8,192 scenarios x 64
replicates at 1M agents x 120 days), held out from training together with
their strain-swapped twins, the trained 5-member ensemble predicts the
simulator's ensemble-mean curves within 5% (peak day, peak height and attack
of each strain, median relative error), an untrained ensemble does not, and
a prediction costs a tiny fraction of simulating the scenario's 64
replicates.  The simulation of two held-out
scenarios is also checked against the committed targets, so the training
set is what the engine produces.
"""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "pandemicthrust_amd", "data", "surrogate_ensemble.npz")
HELDOUT = 256
MAX_MEDIAN_ERROR = 0.05
MIN_GAIN = 100.0
SCHEDULE = ((8000, 1e-3), (8000, 3e-4), (8000, 1e-4))


def test_trained_surrogate_gain_over_simulation():
    from pandemicthrust_amd import ensemble as E
    from pandemicthrust_amd import surrogate as S
    X, Y, meta = E.load_training_set(DATA)
    assert len(X) == 8192 and meta["replicates"] == 64 and meta["agents"] == 1_000_000 and meta["days"] == 120
    ntr = len(X) - HELDOUT
    Xtr, Ytr, Xte, Yte = X[:ntr], Y[:ntr], X[ntr:], Y[ntr:]

    untrained = S.Surrogate(members=5, batch=1024, seed=7)
    P0 = untrained.predict(S.normalize(Xte)).mean(axis=0).reshape(HELDOUT, 2, 120, 3)
    untrained.close()
    err0 = [S.curve_errors(P0[:, k, :, 0], Yte[:, k, :, 0]) for k in range(2)]

    sur = S.Surrogate(members=5, batch=1024, seed=1)
    Xa, Ya = S.augment(Xtr, Ytr)
    sur.set_data(S.normalize(Xa), S.targets(Ya))
    for n, lr in SCHEDULE:
        sur.train(n, lr=lr, seed=n)
    t0 = time.perf_counter()
    P = sur.predict(S.normalize(Xte))
    pred_s = time.perf_counter() - t0
    sur.close()
    Pm = P.mean(axis=0).reshape(HELDOUT, 2, 120, 3)
    for strain in range(2):
        err = S.curve_errors(Pm[:, strain, :, 0], Yte[:, strain, :, 0])
        med = {k: float(np.median(v)) for k, v in err.items()}
        med0 = {k: float(np.median(v)) for k, v in err0[strain].items()}
        print("strain %d held-out median relative errors: trained %s, untrained %s" % (
            strain, {k: round(v, 4) for k, v in med.items()}, {k: round(v, 4) for k, v in med0.items()}))
        for k in ("peak_day", "peak_height", "attack"):
            assert med[k] <= MAX_MEDIAN_ERROR, (strain, k, med[k])
            assert med0[k] > 4 * med[k], (strain, k)

    # The method without the surrogate: simulate the scenarios' 64 replicates.
    pick = [ntr, ntr + 1]
    t0 = time.perf_counter()
    Ysim, info = E.run_ensemble(X[pick], 64, 1_000_000, 120, seed0=E.TRAINING_SET_SEED0 + ntr * 64)
    sim_s = (time.perf_counter() - t0) / len(pick)
    assert np.array_equal(Ysim.astype(np.float16), Y[pick].astype(np.float16))
    gain = sim_s / (pred_s / HELDOUT)
    print("simulation %.3f s per scenario, prediction %.6f s per scenario: gain %.0fx" % (
        sim_s, pred_s / HELDOUT, gain))
    assert gain >= MIN_GAIN
