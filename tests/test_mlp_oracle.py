"""The surrogate's CPU oracle (oracle/mlp_oracle.py), pinned by finite
differences and by exact BF16 rounding cases."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402


def _params(seed):
    rng = np.random.default_rng(seed)
    lay, count = M.layout()
    p = np.zeros(count, np.float64)
    for l, (w, b) in enumerate(lay):
        a = np.sqrt(6.0 / (M.DIMS[l] + M.DIMS[l + 1]))
        p[w:w + M.DIMS[l] * M.DIMS[l + 1]] = rng.uniform(-a, a, M.DIMS[l] * M.DIMS[l + 1])
        p[b:b + M.DIMS[l + 1]] = rng.uniform(-0.1, 0.1, M.DIMS[l + 1])
    return p


def test_layout_count():
    _, count = M.layout()
    assert count == 16 * 512 + 512 + 3 * (512 * 512 + 512) + 512 * 720 + 720


def test_gradients_match_finite_differences():
    p = _params(1)
    rng = np.random.default_rng(2)
    X = rng.uniform(0, 1, (8, 16))
    Y = rng.normal(0, 1, (8, 720))
    loss, g = M.loss_and_grads(p, X, Y, bf16_mode=False)
    lay, count = M.layout()
    probe = []
    for l, (w, b) in enumerate(lay):
        probe += list(rng.integers(w, w + M.DIMS[l] * M.DIMS[l + 1], 4)) + list(rng.integers(b, b + M.DIMS[l + 1], 2))
    h = 1e-5
    for i in probe:
        q = p.copy()
        q[i] += h
        lp, _ = M.loss_and_grads(q, X, Y, bf16_mode=False)
        q[i] -= 2 * h
        lm, _ = M.loss_and_grads(q, X, Y, bf16_mode=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_gelu_derivative():
    x = np.linspace(-6, 6, 101)
    fd = (M.gelu(x + 1e-6) - M.gelu(x - 1e-6)) / 2e-6
    assert np.allclose(fd, M.gelu_grad(x), atol=1e-7)
    assert abs(M.gelu(np.array([0.0]))[0]) == 0 and abs(M.gelu(np.array([10.0]))[0] - 10.0) < 1e-12


def test_bf16_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -2.0 - 2 ** -6, -2.0 - 2 ** -7],
                 np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                       # tie -> even (1.0)
    assert r[2] == 1.0 + 4 * 2 ** -8         # tie -> even (1.015625)
    assert r[3] == 1.0 + 2 ** -7             # above the tie -> up
    assert r[4] == -2.0 - 2 ** -6            # representable (step 2^-6 above 2)
    assert r[5] == -2.0                      # tie -> even
