"""MLP surrogate of Monte Carlo quantities of interest and the control-variate
estimator that uses it.

The network runs on libtsm's BF16 MFMA kernels (FP32 accumulation, FP32
master weights, Adam); this module only standardises inputs and outputs and
drives training.  The surrogate enters the estimate only as a control variate,
so the estimate stays unbiased whatever the network's error:

    mean(Q) ~ mean_p(Q - a g) + a mean_m(g)

with the solver's Q and the surrogate's g on the same n_p paired inputs and g
alone on m >> n_p fresh inputs; a per QoI is fitted on an independent set of
pairs (or passed in), so it is a constant with respect to the paired sample.
The second moment is estimated the same way with Q^2 and g^2, giving the
variance.  The reference has no counterpart (it averages plain samples,
src/fv.cc:595-694).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import TsmError, lib


def _mlib():
    L = lib()
    if not getattr(L, "_mlp_bound", False):
        L.tsm_mlp_last_error.restype = C.c_char_p
        L.tsm_mlp_create.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_ulonglong,
                                     C.POINTER(C.c_void_p)]
        L.tsm_mlp_destroy.argtypes = [C.c_void_p]
        L.tsm_mlp_destroy.restype = None
        L.tsm_mlp_n_params.argtypes = [C.c_void_p]
        L.tsm_mlp_n_params.restype = C.c_long
        L.tsm_mlp_get_params.argtypes = [C.c_void_p, C.c_void_p]
        L.tsm_mlp_set_params.argtypes = [C.c_void_p, C.c_void_p]
        L.tsm_mlp_forward.argtypes = [C.c_void_p, C.c_long, C.c_void_p, C.c_void_p]
        L.tsm_mlp_gradient.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                       C.POINTER(C.c_double), C.c_void_p]
        L.tsm_mlp_train.argtypes = [C.c_void_p, C.c_long, C.c_void_p, C.c_void_p, C.c_int,
                                    C.c_int, C.c_float, C.c_ulonglong, C.c_void_p]
        L.tsm_mlp_gemm_test.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                        C.c_void_p, C.c_int, C.c_void_p]
        L._mlp_bound = True
    return L


def _chk(rc):
    if rc != 0:
        raise TsmError(f"libtsm mlp error {rc}: {_mlib().tsm_mlp_last_error().decode()}")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class MLP:
    """widths[0] -> ... -> widths[-1], GELU between layers, on the GPU."""

    def __init__(self, widths, max_batch=4096, seed=0):
        L = _mlib()
        self.widths = list(widths)
        w = np.asarray(self.widths, dtype=np.int32)
        h = C.c_void_p()
        _chk(L.tsm_mlp_create(len(w) - 1, _p(w), max_batch, seed, C.byref(h)))
        self._h = h
        self.max_batch = max_batch
        self.n_params = int(L.tsm_mlp_n_params(h))

    def close(self):
        if getattr(self, "_h", None):
            _mlib().tsm_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        p = np.empty(self.n_params, np.float32)
        _chk(_mlib().tsm_mlp_get_params(self._h, _p(p)))
        return p

    @params.setter
    def params(self, p):
        p = _f32(p)
        assert p.size == self.n_params
        _chk(_mlib().tsm_mlp_set_params(self._h, _p(p)))

    def __call__(self, x):
        x = _f32(x)
        y = np.empty((x.shape[0], self.widths[-1]), np.float32)
        _chk(_mlib().tsm_mlp_forward(self._h, x.shape[0], _p(x), _p(y)))
        return y

    def gradient(self, x, t):
        x, t = _f32(x), _f32(t)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_double()
        _chk(_mlib().tsm_mlp_gradient(self._h, x.shape[0], _p(x), _p(t), C.byref(loss), _p(g)))
        return loss.value, g

    def train(self, x, t, epochs, batch=256, lr=1e-3, seed=0):
        x, t = _f32(x), _f32(t)
        losses = np.zeros(epochs)
        _chk(_mlib().tsm_mlp_train(self._h, x.shape[0], _p(x), _p(t), epochs, batch, lr, seed,
                                   _p(losses)))
        return losses


def gemm_test(A_bits, B_bits, a_kmajor=True, b_kmajor=True):
    """C = A B through the MLP's MFMA GEMM; A_bits/B_bits are bf16 bit patterns
    stored as (M,K)/(K,M) and (N,K)/(K,N) by the layout flags."""
    A = np.ascontiguousarray(A_bits, np.uint16)
    B = np.ascontiguousarray(B_bits, np.uint16)
    M, K = A.shape if a_kmajor else A.shape[::-1]
    N = B.shape[0] if b_kmajor else B.shape[1]
    out = np.empty((M, N), np.float32)
    _chk(_mlib().tsm_mlp_gemm_test(M, N, K, _p(A), int(a_kmajor), _p(B), int(b_kmajor), _p(out)))
    return out


class Surrogate:
    """Standardised inputs/outputs around an MLP: u in [0,1]^d -> QoI vector."""

    def __init__(self, n_in, n_out, hidden=(256, 256, 256), max_batch=4096, seed=0):
        self.net = MLP([n_in, *hidden, n_out], max_batch=max_batch, seed=seed)
        self.mu_x = np.zeros(n_in)
        self.sd_x = np.ones(n_in)
        self.mu_y = np.zeros(n_out)
        self.sd_y = np.ones(n_out)

    def fit(self, u, Q, epochs=200, batch=128, lr=2e-3, seed=0):
        u, Q = np.asarray(u, np.float64), np.asarray(Q, np.float64)
        self.mu_x, self.sd_x = u.mean(0), u.std(0) + 1e-12
        self.mu_y, self.sd_y = Q.mean(0), Q.std(0) + 1e-30
        return self.net.train((u - self.mu_x) / self.sd_x, (Q - self.mu_y) / self.sd_y, epochs,
                              batch=batch, lr=lr, seed=seed)

    def __call__(self, u):
        u = np.asarray(u, np.float64)
        y = self.net((u - self.mu_x) / self.sd_x).astype(np.float64)
        return y * self.sd_y + self.mu_y


@dataclass
class CVEstimate:
    mean: np.ndarray
    var: np.ndarray
    alpha: np.ndarray          # control coefficient per QoI (first moment)
    mean_se: np.ndarray        # standard error of the mean estimate
    plain_mean_se: np.ndarray  # standard error of the plain mean of the paired samples
    variance_reduction: np.ndarray


def _alpha(q, g):
    """Optimal control coefficient Cov(q, g) / Var(g) per column."""
    qc, gc = q - q.mean(0), g - g.mean(0)
    vg = (gc * gc).sum(0)
    return np.where(vg > 0, (qc * gc).sum(0) / np.where(vg > 0, vg, 1.0), 0.0)


def control_variate(Q_p, G_p, G_m, alpha=None, alpha2=None):
    """Control-variate estimate of E[Q] and Var[Q] per column.

    Q_p, G_p: solver and surrogate QoIs on the same n_p inputs; G_m: surrogate
    QoIs on m independent inputs.  alpha/alpha2: control coefficients of the
    first and second moments, fitted on independent pairs for an unbiased
    estimate (fitted on Q_p, G_p themselves when None: bias O(1/n_p))."""
    Q_p, G_p, G_m = (np.asarray(a, np.float64) for a in (Q_p, G_p, G_m))
    n = Q_p.shape[0]
    a1 = _alpha(Q_p, G_p) if alpha is None else np.asarray(alpha, np.float64)
    a2 = _alpha(Q_p ** 2, G_p ** 2) if alpha2 is None else np.asarray(alpha2, np.float64)
    d1 = Q_p - a1 * G_p
    m1 = d1.mean(0) + a1 * G_m.mean(0)
    m2 = (Q_p ** 2 - a2 * G_p ** 2).mean(0) + a2 * (G_m ** 2).mean(0)
    se = np.sqrt(d1.var(0, ddof=1) / n + a1 ** 2 * G_m.var(0, ddof=1) / G_m.shape[0])
    plain = np.sqrt(Q_p.var(0, ddof=1) / n)
    vr = np.where(se > 0, (plain / np.where(se > 0, se, 1.0)) ** 2, np.inf)
    return CVEstimate(m1, m2 - m1 * m1, a1, se, plain, vr)
