"""Surrogate MLP on the GPU (BF16 MFMA, FP32 accumulation) against the FP64
CPU oracle with the same bf16 rounding points.

Tolerances: GEMM of bf16 inputs (exact products, FP32 sums) rel 1e-5;
forward rel-L2 5e-3 and gradient rel-L2 3e-2 per parameter block (FP32
accumulation order and the bf16 rounding flips it causes)."""
import numpy as np
import pytest

import mlp_oracle as MO
from tensum_amd.surrogate import MLP, Surrogate, gemm_test

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ak,bk", [(True, True), (True, False), (False, True), (False, False)])
def test_mfma_gemm_layouts(ak, bk):
    rng = np.random.default_rng(3)
    M, N, K = 70, 45, 37  # partial tiles in every dimension
    A, B = rng.standard_normal((M, K)), rng.standard_normal((K, N))
    Ab, Bb = MO.bf16(A), MO.bf16(B)
    ref = Ab @ Bb
    As = MO.bf16_bits(A if ak else A.T)
    Bs = MO.bf16_bits(B.T if bk else B)
    C = gemm_test(As, Bs, ak, bk)
    assert np.allclose(C, ref, rtol=1e-5, atol=1e-5 * np.abs(ref).max())


WIDTHS = [6, 256, 256, 256, 66]


def _blocks(widths, flat):
    out, o = [], 0
    for l in range(len(widths) - 1):
        nw = widths[l] * widths[l + 1]
        out.append(flat[o:o + nw]); o += nw
        out.append(flat[o:o + widths[l + 1]]); o += widths[l + 1]
    return out


def test_forward_matches_oracle():
    rng = np.random.default_rng(4)
    net = MLP(WIDTHS, max_batch=512, seed=7)
    p = net.params.astype(np.float64)
    p[:] += 0.01 * rng.standard_normal(p.size)  # non-zero biases
    net.params = p
    x = rng.standard_normal((700, 6))  # two chunks of max_batch
    y = net(x)
    ref, _ = MO.forward(net.params.astype(np.float64), WIDTHS, x, bf=True)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 5e-3


def test_gradient_matches_oracle():
    rng = np.random.default_rng(5)
    net = MLP(WIDTHS, max_batch=256, seed=8)
    x, t = rng.standard_normal((200, 6)), rng.standard_normal((200, 66))
    loss, g = net.gradient(x, t)
    lo, go = MO.loss_and_grad(net.params.astype(np.float64), WIDTHS, x, t, bf=True)
    assert abs(loss - lo) <= 1e-4 * lo
    for a, b in zip(_blocks(WIDTHS, g), _blocks(WIDTHS, go)):
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 3e-2


def _target(u):
    k = np.arange(1, 67)[None, :]
    return np.sin(np.pi * (u @ np.linspace(0.2, 1.0, 6))[:, None] * k / 33.0) + u[:, :1] ** 2


def test_training_reduces_loss_and_generalises():
    rng = np.random.default_rng(6)
    u = rng.random((2048, 6))
    s = Surrogate(6, 66, max_batch=2048, seed=1)
    losses = s.fit(u, _target(u), epochs=150, batch=128, lr=2e-3)
    assert losses[-1] < 0.05 * losses[0]
    ut = rng.random((512, 6))
    err = np.linalg.norm(s(ut) - _target(ut)) / np.linalg.norm(_target(ut) - _target(ut).mean(0))
    assert err < 0.2
