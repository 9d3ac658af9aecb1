"""Pins of the oracle's pointwise functions (CPU only).

Each check is fixed by mathematics, not by retyping the oracle's formula:
closed forms, Tadmor's entropy-conservation identity, consistency with the
exact Euler flux, eigen-structure limits of the TeCNO dissipation, and
sign-pattern known answers for the limiters.  Reference functions:
src/material.h:24-42 (logavg), src/kepec.cc, src/roe_ec.cc, src/kep.cc,
src/simple_avg.cc, src/kepes_tecno_roe.cc, src/kepes_tecno_rusanov.cc,
src/reconstruct.cc, src/limiter.h.
"""
from decimal import Decimal, getcontext

import numpy as np
import pytest

import tensum_amd as T
import oracle_binding as O

GAM = 1.4
ZERO8 = np.zeros(8)


def prm(**kw):
    base = dict(flux="kepec", reconstruct="first")
    base.update(kw)
    return T.Params(**base)


def euler_flux(W, n, gam=GAM):
    Tm, u, v, p = W
    rho = p / Tm
    vn = u * n[0] + v * n[1]
    E = p / (gam - 1) + 0.5 * rho * (u * u + v * v)
    return np.array([rho * vn, rho * u * vn + p * n[0], rho * v * vn + p * n[1], (E + p) * vn])


def ent_vars(W, gam=GAM):
    """Entropy variables of U = -rho s/(gamma-1), s = ln p - gamma ln rho."""
    Tm, u, v, p = W
    rho = p / Tm
    s = np.log(p) - gam * np.log(rho)
    return np.array([(gam - s) / (gam - 1) - rho * (u * u + v * v) / (2 * p), rho * u / p,
                     rho * v / p, -rho / p])


def cons(W, gam=GAM):
    Tm, u, v, p = W
    rho = p / Tm
    return np.array([rho, rho * u, rho * v, p / (gam - 1) + 0.5 * rho * (u * u + v * v)])


def flux(p, Wl, Wr, n, xl=(0, 0), xr=(1, 0), gl=ZERO8, gr=ZERO8):
    return O.num_flux(p, np.array(xl, float), np.array(xr, float), np.array(Wl, float),
                      np.array(Wr, float), gl, gr, np.array(n, float))


STATES = [((1.0, 0.3, -0.2, 1.0), (1.3, 0.1, 0.4, 0.7)),
          ((0.2, 1.206, 1.206, 0.029), (0.56, 0.0, 1.206, 0.3)),
          ((2.0, -0.5, 0.0, 3.0), (2.0, -0.5, 0.0, 3.0000001))]


# ---- logavg ---------------------------------------------------------------
def _logavg_exact(a, b):
    getcontext().prec = 60
    A, B = Decimal(a), Decimal(b)
    if A == B:
        return float(A)
    return float((A - B) / (A.ln() - B.ln()))


@pytest.mark.parametrize("a,b", [(1.0, 3.0), (0.138, 1.5), (1.0, 1.05), (1.0, 1.0 + 1e-9),
                                 (5.0, 4.999), (2.5, 2.5), (1e-3, 1e3)])
def test_logavg_closed_form(a, b):
    got = O.logavg(a, b)
    assert got == pytest.approx(_logavg_exact(a, b), rel=2e-15)
    assert O.logavg(b, a) == pytest.approx(got, rel=4e-16)


def test_logavg_limit_equal_arguments_is_exact():
    for a in (1e-6, 0.7, 1.0, 123.456):
        assert O.logavg(a, a) == a


# ---- two-point fluxes -------------------------------------------------------
@pytest.mark.parametrize("name", ["kepec", "kep", "roe_ec", "simple_avg", "kepes_tecno_roe",
                                  "kepes_tecno_rusanov"])
def test_consistency_with_euler_flux(name):
    W = (1.2, 0.4, -0.3, 0.9)
    n = (0.3, -0.7)
    F = flux(prm(flux=name), W, W, n)
    np.testing.assert_allclose(F, euler_flux(W, n), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("name", ["kepec", "roe_ec"])
@pytest.mark.parametrize("Wl,Wr", STATES[:2])
def test_tadmor_entropy_conservation(name, Wl, Wr):
    n = np.array([0.6, 0.8])
    F = flux(prm(flux=name), Wl, Wr, n)
    dV = ent_vars(Wr) - ent_vars(Wl)
    psi = lambda W: (W[3] / W[0]) * (W[1] * n[0] + W[2] * n[1])  # rho u.n
    lhs, rhs = dV @ F, psi(Wr) - psi(Wl)
    # exact in exact arithmetic; logavg's series stops at u^3 (u < 1e-2), a
    # relative error of at most u^4/9 ~ 1.1e-9 in the averaged quantities
    assert lhs == pytest.approx(rhs, rel=2e-9, abs=1e-13)


def test_kep_flux_is_not_entropy_conservative_but_symmetric():
    Wl, Wr = STATES[0]
    n = (0.6, 0.8)
    a = flux(prm(flux="kep"), Wl, Wr, n)
    b = flux(prm(flux="kep"), Wr, Wl, n)
    np.testing.assert_allclose(a, b, rtol=1e-15)
    # KEP momentum flux: arithmetic means rho_bar u_bar u_bar.n + p_bar n
    rl, rr = Wl[3] / Wl[0], Wr[3] / Wr[0]
    ub = 0.5 * (np.array(Wl[1:3]) + np.array(Wr[1:3]))
    mass = 0.5 * (rl + rr) * (ub @ n)
    np.testing.assert_allclose(a[:3], [mass, *(mass * ub + 0.5 * (Wl[3] + Wr[3]) * np.array(n))],
                               rtol=1e-14)


@pytest.mark.parametrize("name", ["kepes_tecno_roe", "kepes_tecno_rusanov"])
@pytest.mark.parametrize("Wl,Wr", STATES)
def test_tecno_first_order_dissipates_entropy(name, Wl, Wr):
    n = (0.6, 0.8)
    F = flux(prm(flux=name, reconstruct="first"), Wl, Wr, n)
    Fc = flux(prm(flux="kepec"), Wl, Wr, n)
    dV = ent_vars(Wr) - ent_vars(Wl)
    assert dV @ (F - Fc) <= 1e-15


def _jac(W, n, h=1e-7):
    U0 = cons(W)

    def prim(U):
        rho, mx, my, E = U
        u, v = mx / rho, my / rho
        p = (GAM - 1) * (E - 0.5 * rho * (u * u + v * v))
        return (p / rho, u, v, p)

    J = np.zeros((4, 4))
    for k in range(4):
        d = np.zeros(4)
        d[k] = h * max(1.0, abs(U0[k]))
        J[:, k] = (euler_flux(prim(U0 + d), n) - euler_flux(prim(U0 - d), n)) / (2 * d[k])
    return J


@pytest.mark.parametrize("name", ["kepes_tecno_roe", "kepes_tecno_rusanov"])
def test_tecno_dissipation_matches_upwind_limit(name):
    """For a small jump the TeCNO dissipation -(1/2) R Lam R^T dV tends to
    -(1/2)|A_n| dU (Roe) or -(1/2) lam_max dU (Rusanov): pins the scaled
    eigenvectors (R R^T = dU/dV) and the eigenvalues."""
    W = np.array([1.1, 0.35, -0.2, 0.8])
    n = np.array([0.6, 0.8])
    eps = 1e-6
    dW = eps * np.array([0.3, -0.7, 0.5, 0.9])
    Wl, Wr = W - dW, W + dW
    D = -2 * (flux(prm(flux=name), Wl, Wr, n) - flux(prm(flux="kepec"), Wl, Wr, n))
    dU = cons(Wr) - cons(Wl)
    if name == "kepes_tecno_rusanov":
        c = np.sqrt(GAM * W[0])
        lam = abs(W[1:3] @ n) + c
        ref = lam * dU
    else:
        A = _jac(W, n)
        ev, R = np.linalg.eig(A)
        absA = (R @ np.diag(np.abs(ev)) @ np.linalg.inv(R)).real
        ref = absA @ dU
    np.testing.assert_allclose(D, ref, rtol=2e-5, atol=1e-12)


def test_roe_liou_fix_raises_shear_eigenvalues():
    # subsonic normal velocity with a tangential jump: max(|vn|, a) = a > |vn|
    Wl, Wr = STATES[0]
    n = (0.6, 0.8)
    p0 = prm(flux="kepes_tecno_roe")
    a = flux(p0, Wl, Wr, n)
    b = O.num_flux(p0, np.zeros(2), np.ones(2), np.array(Wl), np.array(Wr), ZERO8, ZERO8,
                   np.array(n), ssw=2.0)
    Fc = flux(prm(flux="kepec"), Wl, Wr, n)
    dV = ent_vars(np.array(Wr)) - ent_vars(np.array(Wl))
    assert dV @ (b - Fc) < dV @ (a - Fc) <= 0


def test_flux_scales_with_normal_length():
    Wl, Wr = STATES[0]
    for name in ("kepec", "kepes_tecno_roe", "roe_ec"):
        a = flux(prm(flux=name), Wl, Wr, (0.6, 0.8))
        b = flux(prm(flux=name), Wl, Wr, (1.2, 1.6))
        np.testing.assert_allclose(b, 2 * a, rtol=1e-14)


# ---- limiters and reconstruction -------------------------------------------
@pytest.mark.parametrize("a,b,mm,ma", [(1.0, 2.0, 1.0, 1.0), (-1.0, -3.0, -1.0, -1.0),
                                       (1.0, -1.0, 0.0, -1.0), (0.0, 2.0, 0.0, 0.0),
                                       (-2.0, 0.5, 0.0, 0.5)])
def test_minmod_minabs_sign_patterns(a, b, mm, ma):
    assert O.limited_slope(prm(reconstruct="minmod"), a, b) == 0.5 * mm
    assert O.limited_slope(prm(reconstruct="eno2"), a, b) == 0.5 * ma


def test_tvb_minmod_switch():
    p = prm(reconstruct="tvb_minmod", tvb_M=4.0)
    # |(a+b)/2| = 0.5 < M h^2 = 1 -> central average
    assert O.limited_slope(p, 1.0, 0.0, 0.5) == 0.5 * 0.5
    # h small -> minmod
    assert O.limited_slope(p, 1.0, 0.0, 0.1) == 0.0


def test_van_albada_closed_form():
    p = prm(reconstruct="van_albada")
    a, b = 0.3, 0.7
    s = (2 * a * b + 1e-14) / (a * a + b * b + 1e-14)
    K = 1 / 3
    assert O.limited_slope(p, a, b) == pytest.approx(0.25 * s * ((1 - K * s) * a + (1 + K * s) * b),
                                                     rel=1e-15)
    assert O.limited_slope(p, 1.0, -1.0) == 0.0   # opposite signs -> s = 0


@pytest.mark.parametrize("rec", ["second", "second_muscl", "van_albada", "minmod", "tvb_minmod",
                                 "eno2"])
def test_linear_data_reconstructs_to_the_midpoint(rec):
    zl = np.array([1.0, -2.0, 0.5, 3.0])
    zr = np.array([1.5, -1.0, 0.25, 2.0])
    d = zr - zl
    rl, rr = O.reconstruct(prm(reconstruct=rec, tvb_M=1.0), 1.0, zl, zr, d, d)
    np.testing.assert_allclose(rl, 0.5 * (zl + zr), rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(rr, 0.5 * (zl + zr), rtol=1e-15, atol=1e-15)


def test_first_order_keeps_states_and_minmod_clips_extrema():
    zl, zr = np.array([1.0, 0, 0, 0]), np.array([2.0, 0, 0, 0])
    g = np.array([5.0, 0, 0, 0])  # steep left slope: minmod(2g-d, d)/2 = d/2
    rl, rr = O.reconstruct(prm(reconstruct="first"), 1.0, zl, zr, g, g)
    assert np.array_equal(rl, zl) and np.array_equal(rr, zr)
    rl, rr = O.reconstruct(prm(reconstruct="minmod"), 1.0, zl, zr, g, -g)
    assert rl[0] == 1.5 and rr[0] == 2.0   # right side: minmod(-11, 1) = 0


def test_moments_closed_form_streams():
    """Welford moments of the reference's increment_moments on streams whose
    moments are known: x_s = a + b*s gives mean a + b(S-1)/2,
    M2 = b^2 S(S^2-1)/12, S2 = sum of squares; a constant stream gives M2 = 0."""
    S, nv = 37, 3
    a = np.array([1.0, 0.5, -0.25, 2.0])
    b = np.array([0.01, -0.02, 0.03, 0.005])
    prims = np.array([[a + b * s] * nv for s in range(S)])
    R = np.full(S, 1.0)
    m, q, z = O.moments(prims, np.full(S, 1.4), R)
    s = np.arange(S)
    for k in range(4):
        xs = a[k] + b[k] * s
        assert np.allclose(m[:, k], a[k] + b[k] * (S - 1) / 2, rtol=1e-13, atol=1e-15)
        assert np.allclose(q[:, k], b[k] ** 2 * S * (S * S - 1) / 12, rtol=1e-10)
        assert np.allclose(z[:, k], np.sum(xs * xs), rtol=1e-13)
    rho = prims[:, 0, 3] / (R * prims[:, 0, 0])
    assert np.allclose(m[:, 4], rho.mean(), rtol=1e-13)
    assert np.allclose(q[:, 4], np.sum((rho - rho.mean()) ** 2), rtol=1e-9)
    mc, qc, _ = O.moments(np.ones((5, 2, 4)), np.full(5, 1.4), np.ones(5))
    assert np.all(qc == 0.0) and np.all(mc == 1.0)
