"""Sample inputs: Sobol points and their Owen-type scrambling (host code of libtsm).

Pins: the unscrambled points equal scipy.stats.qmc.Sobol (the same Joe-Kuo
direction numbers, Gray-code order) bit for bit; scrambled points keep the net
properties of the sequence (every dyadic interval of the first 2^m points holds
one point per coordinate; dimensions 1-2 form a (0,m,2)-net); scrambled
estimates of known integrals are unbiased over seeds.
"""
import numpy as np
import pytest
from scipy.stats import qmc

import tensum_amd as T


def test_unscrambled_equals_scipy():
    n, d = 1 << 12, T.lib().tsm_sobol_max_dim()
    assert np.array_equal(T.sobol(n, d, scramble=False), qmc.Sobol(d, scramble=False).random(n))


def test_skip_continues_the_sequence():
    a = T.sobol(64, 5, scramble=True, seed=9)
    b = T.sobol(32, 5, scramble=True, seed=9, skip=32)
    assert np.array_equal(a[32:], b)


@pytest.mark.parametrize("seed", [0, 1, 12345])
def test_scrambled_points_are_nets(seed):
    m = 10
    x = T.sobol(1 << m, 6, seed=seed)
    assert np.all((x > 0) & (x < 1))
    for j in range(6):  # one point per interval [k/2^m, (k+1)/2^m)
        cells = np.floor(x[:, j] * (1 << m)).astype(int)
        assert np.array_equal(np.sort(cells), np.arange(1 << m))
    for a in range(m + 1):  # (0,m,2)-net: boxes 2^-a x 2^-(m-a) hold one point each
        ca = np.floor(x[:, 0] * (1 << a)).astype(int)
        cb = np.floor(x[:, 1] * (1 << (m - a))).astype(int)
        assert np.unique(ca * (1 << (m - a)) + cb).size == 1 << m


def test_seeds_give_different_points():
    a, b = T.sobol(256, 4, seed=1), T.sobol(256, 4, seed=2)
    assert not np.any(a == b)


def test_scrambled_estimates_are_unbiased():
    """E over seeds of the QMC mean of f(x) = prod_j 2 x_j (integral 1) is 1;
    the spread is far below plain MC's (randomised QMC on a smooth integrand)."""
    d, n = 5, 256
    est = np.array([np.prod(2 * T.sobol(n, d, seed=s), axis=1).mean() for s in range(200)])
    se = est.std(ddof=1) / np.sqrt(est.size)
    assert abs(est.mean() - 1.0) < 4 * se + 1e-12
    mc_sd = np.sqrt((4.0 / 3.0) ** d - 1) / np.sqrt(n)  # sd of a plain MC mean of n points
    assert est.std(ddof=1) < 0.2 * mc_sd


def test_argument_errors():
    with pytest.raises(T.TsmError):
        T.sobol(4, 17)
    with pytest.raises(T.TsmError):
        T.sobol(4, 0)
