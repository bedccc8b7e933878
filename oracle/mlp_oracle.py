"""CPU ORACLE — TEST INFRASTRUCTURE ONLY.

numpy restatement of the surrogate MLP of include/pandemic/pt_surrogate.h
(16 -> 512 x 4 (GELU) -> 720, mean-squared error) for checking the GPU
kernels.  With bf16=True every value the GPU stores in BF16 (weights,
activations, output gradients fed to the matrix products) is rounded to
BF16 (round to nearest even) at the same point; products accumulate in
float64.  With bf16=False the whole computation is float64.  GELU is the
exact x * Phi(x).
"""
import math

import numpy as np

DIMS = [16, 512, 512, 512, 512, 720]
LAYERS = len(DIMS) - 1


def layout():
    """[(w_offset, b_offset)] per layer and the parameter count of one member."""
    off, out = 0, []
    for l in range(LAYERS):
        w = off
        off += DIMS[l] * DIMS[l + 1]
        out.append((w, off))
        off += DIMS[l + 1]
    return out, off


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest, ties to even); returns float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


_erf = np.vectorize(math.erf)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    cdf = 0.5 * (1.0 + _erf(x / math.sqrt(2.0)))
    pdf = np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)
    return cdf + x * pdf


def unpack(params):
    lay, _ = layout()
    Ws, bs = [], []
    for l, (w, b) in enumerate(lay):
        Ws.append(params[w:w + DIMS[l] * DIMS[l + 1]].reshape(DIMS[l], DIMS[l + 1]))
        bs.append(params[b:b + DIMS[l + 1]])
    return Ws, bs


def forward(params, X, bf16_mode=True):
    """Predictions and the values backward needs."""
    rnd = (lambda a: bf16(a).astype(np.float64)) if bf16_mode else (lambda a: np.asarray(a, np.float64))
    Ws, bs = unpack(params)
    acts, pres = [rnd(X)], []
    for l in range(LAYERS):
        z = acts[-1] @ rnd(Ws[l]).astype(np.float64) + bs[l].astype(np.float64)
        if bf16_mode:
            z = z.astype(np.float32).astype(np.float64)
        if l < LAYERS - 1:
            pres.append(z)
            acts.append(rnd(gelu(z).astype(np.float32)) if bf16_mode else gelu(z))
        else:
            pred = z
    return pred, acts, pres


def loss_and_grads(params, X, Y, bf16_mode=True):
    """(loss, gradient) of the mean-squared error, laid out like params."""
    rnd = (lambda a: bf16(a).astype(np.float64)) if bf16_mode else (lambda a: np.asarray(a, np.float64))
    Ws, _ = unpack(params)
    pred, acts, pres = forward(params, X, bf16_mode)
    n = pred.size
    d = pred - np.asarray(Y, np.float64)
    loss = float((d * d).sum() / n)
    dz = 2.0 * d / n
    if bf16_mode:
        dz = dz.astype(np.float32).astype(np.float64)
    lay, count = layout()
    g = np.zeros(count, np.float64)
    for l in range(LAYERS - 1, -1, -1):
        w, b = lay[l]
        g[w:w + DIMS[l] * DIMS[l + 1]] = (acts[l].T @ rnd(dz)).ravel()
        g[b:b + DIMS[l + 1]] = dz.sum(axis=0)
        if l == 0:
            break
        da = rnd(dz) @ rnd(Ws[l]).T
        dz = da * gelu_grad(pres[l - 1])
        if bf16_mode:
            dz = dz.astype(np.float32).astype(np.float64)
    return loss, g

