"""CPU oracle of the surrogate MLP: forward and backward in FP64 (numpy).

Test infrastructure only: imported by tests/ as the checker of the MFMA
kernels, never by the product.  The network has no counterpart in the
reference.  Parameters are flat: W_0 (n_1 x n_0, row-major), b_0, W_1, b_1, ...

With bf16=True the rounding points of the GPU path are reproduced (inputs,
weights, hidden activations and backward signals rounded to bfloat16, round to
nearest even) while every product accumulates in FP64, so the GPU result
differs from it only by FP32 accumulation and the rare bf16 rounding flip that
accumulation causes.  With bf16=False it is the exact FP64 network, whose
gradient the tests check against finite differences.
"""
from __future__ import annotations

import math

import numpy as np

_SQRT1_2 = 1.0 / math.sqrt(2.0)
_erf = np.vectorize(math.erf, otypes=[np.float64])


def bf16(x):
    """Round to the nearest bfloat16 (ties to even); returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(x):
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z * _SQRT1_2))


def gelu_d(z):
    return 0.5 * (1.0 + _erf(z * _SQRT1_2)) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def unpack(params, widths):
    out, o = [], 0
    for l in range(len(widths) - 1):
        n_in, n_out = widths[l], widths[l + 1]
        W = params[o:o + n_in * n_out].reshape(n_out, n_in)
        o += n_in * n_out
        b = params[o:o + n_out]
        o += n_out
        out.append((np.asarray(W, np.float64), np.asarray(b, np.float64)))
    return out


def forward(params, widths, x, bf=True):
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params, widths)
    xs, zs = [r(x)], []
    for l, (W, b) in enumerate(layers):
        z = xs[-1] @ r(W).T + b
        zs.append(z)
        if l + 1 < len(layers):
            xs.append(r(gelu(z)))
    return zs[-1], (xs, zs)


def loss_and_grad(params, widths, x, t, bf=True):
    """Mean squared error over rows and outputs, and its flat gradient."""
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params, widths)
    y, (xs, zs) = forward(params, widths, x, bf)
    n = y.size
    loss = float(np.sum((y - t) ** 2) / n)
    dz = r(2.0 * (y - t) / n)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        dW = dz.T @ xs[l]
        db = dz.sum(axis=0)
        grads[l] = (dW, db)
        if l > 0:
            dz = r((dz @ r(W)) * gelu_d(zs[l - 1]))
    flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
    return loss, flat
