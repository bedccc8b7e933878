"""Surrogate MLP on gfx950 MFMA against the CPU oracle (oracle/mlp_oracle.py).

Tolerances: the GPU rounds weights, activations and output gradients to BF16
and accumulates in FP32; the oracle (bf16 mode) rounds at the same points and
accumulates in float64, so the two differ only where FP32 accumulation or
erff/expf move a value across a BF16 rounding boundary.  Gradients: relative
Frobenius error per layer <= 2e-2 against the BF16 oracle and <= 6e-2
against the all-float64 oracle; loss: relative error <= 1e-3.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


def _data(n, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0, 1, (n, 16)).astype(np.float32)
    W = rng.normal(0, 1.0, (16, 720))
    Y = np.sin(X @ W).astype(np.float32)   # smooth learnable target
    return X, Y


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_gradients_match_oracle():
    from pandemicthrust_amd.surrogate import Surrogate
    s = Surrogate(members=2, batch=64, seed=3)
    X, Y = _data(256)
    s.set_data(X, Y)
    rows = np.arange(64, 128)
    g, loss = s.gradients(rows)
    P = s.params()
    lay, count = M.layout()
    for k in range(2):
        l16, g16 = M.loss_and_grads(P[k].astype(np.float64), X[rows], Y[rows], bf16_mode=True)
        l64, g64 = M.loss_and_grads(P[k].astype(np.float64), X[rows], Y[rows], bf16_mode=False)
        assert abs(loss[k] - l16) <= 1e-3 * l16, (k, loss[k], l16)
        for l, (w, b) in enumerate(lay):
            sl = slice(w, b + M.DIMS[l + 1])
            assert _rel(g[k, sl], g16[sl]) <= 2e-2, (k, l, _rel(g[k, sl], g16[sl]))
            assert _rel(g[k, sl], g64[sl]) <= 6e-2, (k, l, _rel(g[k, sl], g64[sl]))
    s.close()


def test_predictions_match_oracle():
    from pandemicthrust_amd.surrogate import Surrogate
    s = Surrogate(members=3, batch=64, seed=5)
    X, _ = _data(100, seed=1)
    P = s.params()
    out = s.predict(X)          # 100 rows: two chunks, the second partial
    assert out.shape == (3, 100, 720)
    for k in range(3):
        ref, _, _ = M.forward(P[k].astype(np.float64), X, bf16_mode=True)
        assert _rel(out[k], ref) <= 1e-2, (k, _rel(out[k], ref))
    assert not np.allclose(out[0], out[1])      # members start from different weights
    s.close()


def test_adam_step_matches_closed_form():
    from pandemicthrust_amd.surrogate import Surrogate
    s = Surrogate(members=2, batch=64, seed=7)
    X, Y = _data(64, seed=2)
    s.set_data(X, Y)
    p0 = s.params().astype(np.float64)
    g, _ = s.gradients(np.arange(64))
    lr = 1e-3
    s.train(1, lr=lr, seed=0)   # one step over a permutation of the same 64 rows
    p1 = s.params().astype(np.float64)
    # First Adam step: m = 0.1 g, v = 0.001 g^2, bias-corrected m/v^0.5 = g/|g|.
    g64 = g.astype(np.float64)
    m, v = 0.1 * g64, 0.001 * g64 * g64
    expect = p0 - lr * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    # The batch is the same set of rows in another order, so FP32 sums may
    # differ slightly; where a gradient is near zero its sign (and so the
    # step's sign) is not determined by the data, so those are left out.
    sure = np.abs(g64) > 1e-6
    assert sure.mean() > 0.5
    assert np.max(np.abs(p1 - expect)[sure]) <= 2e-5 + 1e-3 * lr, np.max(np.abs(p1 - expect)[sure])
    assert (np.abs(p1 - p0) <= lr * (1 + 1e-3) + 1e-7).all()     # an Adam first step moves at most lr
    s.close()


def test_training_reduces_the_loss():
    from pandemicthrust_amd.surrogate import Surrogate
    s = Surrogate(members=5, batch=256, seed=11)
    X, Y = _data(2048, seed=3)
    s.set_data(X, Y)
    loss = s.train(300, lr=1e-3, seed=1)
    assert loss.shape == (300, 5)
    first, last = loss[:10].mean(axis=0), loss[-10:].mean(axis=0)
    assert (last < 0.5 * first).all(), (first, last)
    s.close()


def test_graph_replayed_training_equals_eager_training():
    """Training steps replayed from the captured HIP graph give the same
    parameters and losses as the same steps enqueued kernel by kernel: one
    call of 7 steps (7 replays), then 3 calls of 2 steps with host work and
    synchronisation between them, each compared with the eager run."""
    from pandemicthrust_amd.surrogate import Surrogate
    X, Y = _data(1024, seed=5)
    runs = []
    for graphs in (True, False):
        s = Surrogate(members=3, batch=256, seed=4)
        s.use_graphs(graphs)
        s.set_data(X, Y)
        losses = [s.train(7, lr=1e-3, seed=1)]
        params = [s.params()]
        for k in range(3):
            losses.append(s.train(2, lr=3e-4, seed=10 + k))
            params.append(s.params())
            np.asarray(params[-1]).sum()    # host work between the calls
        s.close()
        runs.append((losses, params))
    for a, b in zip(runs[0][0], runs[1][0]):
        assert np.array_equal(a, b)
    for a, b in zip(runs[0][1], runs[1][1]):
        assert np.array_equal(a, b)
