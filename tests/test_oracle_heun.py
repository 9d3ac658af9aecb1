"""Pins of the oracle's Heun scheme (time_scheme heuns, src/fv.cc:1675-1686 as
its coefficients mean it: the reference's own update overwrites the
(1 - r/2) U_r term and gives NaN after one step, so nothing of its output
can be matched).  One step is the two-stage SSP scheme
U1 = U0 - dt/A R(U0), U2 = (U0 + U1 - dt/A R(U1)) / 2 built here from the
oracle's residual, and the scheme is second order in time.
"""
import numpy as np

import tensum_amd as T
import oracle_binding as O
from cases import conserved, vortex, vortex_mesh


def _prim(U, gam=1.4, R=1.0):
    rho, mx, my, E = U.T
    u, v = mx / rho, my / rho
    p = (gam - 1) * (E - 0.5 * rho * (u * u + v * v))
    return np.stack([p / (R * rho), u, v, p], 1)


def test_heun_step_is_the_two_stage_ssp_scheme():
    m = vortex_mesh(17)
    prm = T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=10.0,
                   time_scheme="heuns")
    W0 = vortex(m.x, m.y)
    o = O.Oracle(m, prm)
    o.set_state(W0)
    dt = o.compute_dt()
    A = o.geometry()["dcarea"][:, None]
    U0 = conserved(W0)

    def L(W):
        o.set_state(W)
        return -(dt / A) * o.residual()

    U1 = U0 + L(W0)
    U2 = 0.5 * U0 + 0.5 * (U1 + L(_prim(U1)))
    s = O.Oracle(m, prm)
    s.set_state(W0)
    n, log = s.run(1, want_log=True)
    assert n == 1 and log[0, 1] == dt
    np.testing.assert_allclose(s.get_state(), _prim(U2), rtol=1e-12, atol=1e-14)


def test_heun_is_second_order_in_time():
    m = vortex_mesh(17)
    W0 = vortex(m.x, m.y)

    def run(scheme, dt, t=0.4):
        o = O.Oracle(m, T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.0,
                                 time_step=dt, final_time=10.0, time_scheme=scheme))
        o.set_state(W0)
        assert o.run(int(round(t / dt))) == int(round(t / dt))
        return conserved(o.get_state())

    ref = run("rk4", 0.4 / 256)
    errs = [np.abs(run("heuns", 0.4 / k) - ref).max() for k in (16, 32, 64)]
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > 1.8) and np.all(rates < 2.3), (errs, rates)
