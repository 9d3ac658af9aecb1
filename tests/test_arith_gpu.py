"""The stage kernels' quotient and square root against the device's IEEE ones.

fdiv and fsqrt (tensum_amd/csrc/tsm_physics.h) are the refinement steps of
the IEEE division and square-root sequences without their range scaling and
special-value fix-ups.  This sweeps operand exponents far past anything the
solver can hold -- densities, pressures and temperatures down to 1e-150
(near vacuum) and up to 1e150, geometric lengths of any mesh -- with random
mantissas and both signs, and requires the results to be those of `/` and
sqrt() bit for bit there.  Outside that range the two may differ (the range
scaling is what they drop); the test also pins the documented behaviour at
zero: fdiv(0, b) = 0, and fsqrt(0) is NaN where sqrt(0) = 0 (no stage-kernel
square root takes a zero operand from a finite state).
"""
import numpy as np
import pytest

import tensum_amd as T

pytestmark = pytest.mark.gpu


def _sweep(lo, hi, step, n, rng):
    e = np.repeat(np.arange(lo, hi + 1, step), n)
    m = rng.uniform(1.0, 2.0, e.size)
    return np.ldexp(m, e)


def test_fdiv_fsqrt_equal_ieee_over_the_solver_range():
    print("arith: exponent sweep", flush=True)
    rng = np.random.default_rng(11)
    # every exponent pair in [-500, 500]^2 (step 4), both signs
    ea = _sweep(-500, 500, 4, 1, rng)
    a = np.repeat(ea, ea.size) * rng.choice([-1.0, 1.0], ea.size * ea.size)
    b = np.tile(ea, ea.size) * rng.choice([-1.0, 1.0], ea.size * ea.size)
    out = T.check_arith(a, b)
    assert np.array_equal(out[:, 0].view(np.int64), out[:, 1].view(np.int64))
    # square roots over [1e-300, 1e300], every exponent, 64 mantissas each
    x = _sweep(-1000, 1000, 1, 64, rng)
    out = T.check_arith(x, np.ones_like(x))
    assert np.array_equal(out[:, 2].view(np.int64), out[:, 3].view(np.int64))


def test_fdiv_fsqrt_at_zero_and_state_like_operands():
    rng = np.random.default_rng(12)
    # quotients the kernels form: p/(R T), 0.5/(R T), 1/rho, c^2/(gamma-1), ...
    rho = 10.0 ** rng.uniform(-150, 150, 200000)
    T_ = 10.0 ** rng.uniform(-150, 150, 200000)
    out = T.check_arith(rho, T_)
    assert np.array_equal(out[:, 0].view(np.int64), out[:, 1].view(np.int64))
    assert np.array_equal(out[:, 2].view(np.int64), out[:, 3].view(np.int64))
    out = T.check_arith(np.array([0.0, -0.0]), np.array([3.0, 7.0]))
    assert np.array_equal(out[:, 0], out[:, 1])          # fdiv(0, b) = 0
    assert np.all(out[:, 3] == 0.0) and np.all(np.isnan(out[:, 2]))  # fsqrt(0) is NaN
