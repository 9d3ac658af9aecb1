"""Pins of the oracle's surface quantities and of the gradients they use.

 * Skin friction (src/output_skin_friction.cc:43-95): plane Couette flow
   u = U y between a wall at rest and one moving with U has cf = mu U / H on
   both walls (the vertex gradients of a linear field are exact), with
   constant and Sutherland viscosity; the heat flux of its uniform
   temperature is zero.
 * Heat flux (src/output_heatflux.cc:43-90): fluid at rest with
   T = 1/(a + b y) (1/T linear, so the entropy variable l and its vertex
   gradient are exact) has -k dT/dn = -k b/a^2 on the lower wall and
   k b/(a + b)^2 on the upper, k = mu cp / Pr.
 * The gradients dE holds after a step are those of the last stage's input
   state (src/fv.cc:1476, :2295-2304): for Heun's method the forward-Euler
   state U0 - dt/A R(U0), built here from the oracle's residual; before a
   step, those of the current state (src/fv.cc:2238).
"""
import numpy as np

import tensum_amd as T
import oracle_binding as O
from cases import conserved, rel_l2, vortex, vortex_mesh
from test_oracle_heun import _prim
from test_oracle_walls import couette_mesh

GAM, R = 1.4, 1.0


def _ns(**kw):
    return T.Params(**{**dict(model="ns", flux="kepec", reconstruct="first", mu_ref=0.05,
                              prandtl=0.72, cfl=0.5, final_time=10.0), **kw})


def _walls(m):
    bt = np.asarray(m.bface_tag)
    return int((bt == 1).sum()), int((bt == 3).sum())


def test_couette_skin_friction_and_zero_heat_flux():
    U, Tw = 0.4, 1.3
    m = couette_mesh(13, U, Tw)
    w = np.stack([Tw + 0 * m.y, U * m.y, 0 * m.y, 1.0 + 0 * m.y], 1)
    for prm, mu in ((_ns(), 0.05),
                    (_ns(viscosity="sutherland", mu_ref=0.02, T_0=0.4),
                     0.02 * Tw ** 1.5 / (Tw + 0.4))):
        o = O.Oracle(m, prm)
        o.set_state(w)
        nb, nt = _walls(m)
        for tags, n in (([1], nb), ([3], nt), ([1, 3], nb + nt)):
            s = o.surface(tags)
            assert s.shape == (n, 2)
            np.testing.assert_allclose(s[:, 0], mu * U, rtol=1e-12)
            assert np.abs(s[:, 1]).max() < 1e-12


def test_conduction_heat_flux_on_the_walls():
    a, b = 1.0, -0.3                       # T from 1 at y = 0 to 1/0.7 at y = 1
    T0, T1 = 1.0 / a, 1.0 / (a + b)
    m = couette_mesh(13, 0.0, T0)
    bc = dict(m.bc)
    bc[3] = ("noslip", (T1, 0.0, 0.0, 1.0))
    m = T.Mesh(m.x, m.y, m.bface_tag, m.bface_v, m.cell_v, m.periodic_ptr, m.periodic_v, bc)
    w = np.stack([1.0 / (a + b * m.y), 0 * m.y, 0 * m.y, 1.0 + 0 * m.y], 1)
    cp = GAM * R / (GAM - 1)
    for prm in (_ns(), _ns(viscosity="sutherland", mu_ref=0.02, T_0=0.4)):
        o = O.Oracle(m, prm)
        o.set_state(w)
        for tag, Tw, dTdy, ny in ((1, T0, -b / a ** 2, -1.0), (3, T1, -b / (a + b) ** 2, 1.0)):
            mu = 0.05 if prm.viscosity == "constant" else 0.02 * Tw ** 1.5 / (Tw + 0.4)
            k = mu * cp / 0.72
            s = o.surface([tag])
            np.testing.assert_allclose(s[:, 1], -k * dTdy * ny, rtol=1e-11)
            assert np.abs(s[:, 0]).max() < 1e-12


def test_stage_gradients_are_those_of_the_last_stage_input():
    m = vortex_mesh(17)
    prm = T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=10.0,
                   time_scheme="heuns")
    W0 = vortex(m.x, m.y)
    o = O.Oracle(m, prm)
    o.set_state(W0)
    g0 = o.stage_gradients()
    assert np.array_equal(g0, o.gradients())           # before a step: the current state
    dt = o.compute_dt()
    A = o.geometry()["dcarea"][:, None]
    U1 = conserved(W0) - (dt / A) * o.residual()
    W1 = _prim(U1)
    s = O.Oracle(m, prm)
    s.set_state(W0)
    assert s.run(1) == 1
    gs = s.stage_gradients()
    r = O.Oracle(m, prm)
    r.set_state(W1)
    g1 = r.gradients()
    r.set_state(s.get_state())
    g2 = r.gradients()
    assert rel_l2(gs, g1) < 1e-12
    # the three candidates differ by far more than the rounding above
    assert rel_l2(g0, g1) > 1e-4 and rel_l2(g2, g1) > 1e-4
