"""Surrogate pieces that need no GPU: the MLP oracle itself (bf16 rounding,
finite-difference gradient check) and the control-variate estimator
(unbiasedness and variance reduction on a problem with a known mean)."""
import numpy as np

import mlp_oracle as MO
from tensum_amd.surrogate import control_variate


def test_bf16_round_to_nearest_even():
    one = 1.0
    cases = {one: one, one + 2.0 ** -8: one, one + 3 * 2.0 ** -8: one + 2.0 ** -6,
             one + 2.0 ** -7: one + 2.0 ** -7, -2.5: -2.5, 3.0e38: 2.9975e38}
    for x, want in cases.items():
        got = MO.bf16(np.array([x]))[0]
        assert got == np.float32(want) or abs(got - want) <= 2.0 ** -8 * abs(want), (x, got, want)
    assert MO.bf16_bits(np.array([1.0]))[0] == 0x3F80


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    widths = [3, 7, 5, 2]
    n = sum(widths[l] * widths[l + 1] + widths[l + 1] for l in range(3))
    p = rng.standard_normal(n) * 0.5
    x, t = rng.standard_normal((9, 3)), rng.standard_normal((9, 2))
    _, g = MO.loss_and_grad(p, widths, x, t, bf=False)
    h = 1e-6
    fd = np.empty(n)
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        fd[i] = (MO.loss_and_grad(p + e, widths, x, t, bf=False)[0]
                 - MO.loss_and_grad(p - e, widths, x, t, bf=False)[0]) / (2 * h)
    assert np.max(np.abs(fd - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))


def test_oracle_bf16_path_is_close_to_fp64():
    rng = np.random.default_rng(1)
    widths = [6, 32, 32, 4]
    n = sum(widths[l] * widths[l + 1] + widths[l + 1] for l in range(3))
    p = rng.standard_normal(n) * 0.3
    x = rng.standard_normal((64, 6))
    y64, _ = MO.forward(p, widths, x, bf=False)
    y16, _ = MO.forward(p, widths, x, bf=True)
    assert np.linalg.norm(y16 - y64) / np.linalg.norm(y64) < 2e-2


def _toy(u):
    return np.stack([np.sin(2 * np.pi * u[:, 0]) + u[:, 1] ** 2,
                     np.exp(u[:, 0] * u[:, 1])], 1)


TOY_MEAN = np.array([1.0 / 3.0, 1.3179021514544038])  # int sin = 0, int u^2 = 1/3; int int e^{uv}


def test_control_variate_unbiased_and_reduces_variance():
    rng = np.random.default_rng(5)
    def surrogate(u):  # imperfect: biased and noisy, but correlated
        return 0.8 * _toy(u) + 0.1 + 0.05 * np.cos(7 * u[:, :1])
    # control coefficients from an independent pilot set
    up = rng.random((200, 2))
    from tensum_amd.surrogate import _alpha
    a1 = _alpha(_toy(up), surrogate(up))
    a2 = _alpha(_toy(up) ** 2, surrogate(up) ** 2)
    cv, plain = [], []
    for _ in range(300):
        u = rng.random((32, 2))
        um = rng.random((4096, 2))
        e = control_variate(_toy(u), surrogate(u), surrogate(um), alpha=a1, alpha2=a2)
        cv.append(e.mean)
        plain.append(_toy(u).mean(0))
    cv, plain = np.array(cv), np.array(plain)
    se = cv.std(0, ddof=1) / np.sqrt(cv.shape[0])
    assert np.all(np.abs(cv.mean(0) - TOY_MEAN) < 4 * se)
    assert np.all(plain.var(0) / cv.var(0) > 5.0)
