"""Pins of the oracle's steady mode (local time stepping, src/fv.cc:1499-1621,
:2230-2395).  Steady mode shares the residual and the dt formula with the
unsteady path pinned elsewhere; what is new is that every vertex keeps its own
dt_i = cfl * A_i / lambda_i, that the update uses it, and the max_iter /
min_residue stops.  The checks:
 * a fluid at rest has lambda_i = c * (sum of its dual-face normal lengths),
   lengths computed here from the triangles, not from the oracle's face list;
 * the smallest local dt is the unsteady global dt;
 * one steady SSP-RK3 step equals the textbook Shu-Osher stages built here
   from the oracle's residual with the local dt vector;
 * with a fixed time_step the steady run is the unsteady run bit for bit;
 * the loop stops after max_iter steps, or at the first step whose normalised
   residual is <= min_residue.
"""
import numpy as np
import pytest

import tensum_amd as T
import oracle_binding as O
from cases import conserved, riemann_mesh, vortex, vortex_mesh

GAM = 1.4


def _channel(n, seed=3):
    """Unit square: free stream in on the left, out on the right, slip walls."""
    return T.jittered_mesh(n, box=(0.0, 1.0, 0.0, 1.0), jitter=0.2, seed=seed,
                           bc={1: ("slip",), 3: ("slip",), 2: ("outlet",),
                               4: ("farfield", (1.0, 0.5, 0.0, 1.0))})


def _bump(x, y):
    """Free stream (T,u,v,p) = (1, 0.5, 0, 1) with a smooth pressure/temperature bump."""
    g = np.exp(-40.0 * ((x - 0.45) ** 2 + (y - 0.5) ** 2))
    return np.stack([1.0 + 0.1 * g, 0.5 + 0.05 * g, 0.03 * g, 1.0 + 0.2 * g], 1)


def _steady(**kw):
    base = dict(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.5, time_mode="steady",
                max_iter=10 ** 6)
    return T.Params(**{**base, **kw})


def _dual_face_lengths(mesh):
    """Sum over the dual faces around each vertex of |normal|.  A dual face
    runs from the centroid of one cell through the edge midpoint to the
    centroid of the other, so its normal has the length of the chord between
    the two centroids (centroid to midpoint on a boundary edge); a boundary
    edge adds half its length at each end.  Computed from the triangles."""
    x, y = mesh.x, mesh.y
    cells = mesh.cell_v.reshape(-1, 3)
    g = np.stack([x[cells].mean(1), y[cells].mean(1)], 1)
    around = {}
    for c, (a, b, d) in enumerate(cells):
        for i, j in ((a, b), (b, d), (d, a)):
            around.setdefault((min(i, j), max(i, j)), []).append(c)
    tot = np.zeros(mesh.n_vertex)
    for (i, j), cs in around.items():
        if len(cs) == 2:
            ln = np.hypot(*(g[cs[0]] - g[cs[1]]))
        else:
            m = 0.5 * np.array([x[i] + x[j], y[i] + y[j]])
            ln = np.hypot(*(g[cs[0]] - m))
        tot[i] += ln
        tot[j] += ln
    for i, j in mesh.bface_v.reshape(-1, 2):
        h = 0.5 * np.hypot(x[j] - x[i], y[j] - y[i])
        tot[i] += h
        tot[j] += h
    return tot


def test_local_dt_of_a_fluid_at_rest_closed_form():
    mesh = riemann_mesh(9)
    o = O.Oracle(mesh, _steady(cfl=0.7))
    Tm = 1.3
    o.set_state(np.tile([Tm, 0.0, 0.0, 2.0], (mesh.n_vertex, 1)))
    dt_global = o.compute_dt()
    dtl = o.local_dt()
    c = np.sqrt(GAM * 1.0 * Tm)
    expect = 0.7 * o.geometry()["dcarea"] / (c * _dual_face_lengths(mesh))
    np.testing.assert_allclose(dtl, expect, rtol=1e-13)
    assert np.ptp(dtl) > 0.1 * dtl.min()  # jittered cells: the steps really differ
    assert dt_global == dtl.min()


def test_smallest_local_dt_is_the_unsteady_global_dt():
    mesh = _channel(17)
    W = _bump(mesh.x, mesh.y)
    o = O.Oracle(mesh, _steady())
    o.set_state(W)
    o.compute_dt()
    dtl = o.local_dt()
    u = O.Oracle(mesh, T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.5,
                                final_time=10.0))
    u.set_state(W)
    assert u.compute_dt() == dtl.min()
    assert np.all(u.local_dt() == dtl.min())


def _prim(U, R=1.0):
    rho, mx, my, E = U.T
    u, v = mx / rho, my / rho
    p = (GAM - 1) * (E - 0.5 * rho * (u * u + v * v))
    return np.stack([p / (R * rho), u, v, p], 1)


@pytest.mark.parametrize("flux,rec", [("kepes_tecno_roe", "minmod"), ("kepec", "first")])
def test_steady_step_is_shu_osher_with_local_dt(flux, rec):
    mesh = _channel(17)
    W0 = _bump(mesh.x, mesh.y)
    prm = _steady(flux=flux, reconstruct=rec)
    o = O.Oracle(mesh, prm)
    o.set_state(W0)
    o.compute_dt()
    dt = o.local_dt()[:, None]
    A = o.geometry()["dcarea"][:, None]
    U0 = conserved(W0)

    def L(W):
        o.set_state(W)
        return -(dt / A) * o.residual()

    U1 = U0 + L(W0)
    U2 = 0.75 * U0 + 0.25 * (U1 + L(_prim(U1)))
    U3 = U0 / 3.0 + 2.0 / 3.0 * (U2 + L(_prim(U2)))
    s = O.Oracle(mesh, prm)
    s.set_state(W0)
    assert s.run(1) == 1
    np.testing.assert_allclose(s.get_state(), _prim(U3), rtol=1e-12, atol=1e-14)
    # and the step is not the one with the global dt
    g = O.Oracle(mesh, T.Params(flux=flux, reconstruct=rec, cfl=0.5, final_time=10.0))
    g.set_state(W0)
    g.run(1)
    assert np.abs(g.get_state() - s.get_state()).max() > 1e-6


def test_fixed_time_step_steady_equals_unsteady():
    mesh = vortex_mesh(17)
    W = vortex(mesh.x, mesh.y)
    out = []
    for mode in ("steady", "unsteady"):
        o = O.Oracle(mesh, T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.0,
                                    time_step=0.01, time_mode=mode, final_time=10.0))
        o.set_state(W)
        n, log = o.run(12, want_log=True)
        assert n == 12
        out.append((o.get_state(), log))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_steady_loop_stops_at_max_iter():
    mesh = _channel(9)
    o = O.Oracle(mesh, _steady(max_iter=7))
    o.set_state(_bump(mesh.x, mesh.y))
    assert o.run(100) == 7
    assert o.run(100) == 0


def test_steady_loop_stops_at_min_residue_and_converges():
    mesh = _channel(17)
    o = O.Oracle(mesh, _steady(min_residue=1e-4, cfl=0.8))
    o.set_state(_bump(mesh.x, mesh.y))
    n, log = o.run(5000, want_log=True)
    assert 0 < n < 5000
    res = log[:, 3]
    assert res[-1] <= 1e-4 and np.all(res[:-1] > 1e-4)
    assert res[0] == 1.0  # normalised by the first step's norm
    # the bump has left the channel: the state is the free stream
    W = o.get_state()
    np.testing.assert_allclose(W, np.tile([1.0, 0.5, 0.0, 1.0], (mesh.n_vertex, 1)), atol=2e-3)
