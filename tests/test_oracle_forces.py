"""Pins of the oracle's force integrals (src/force.cc:118-217).
 * A fluid at rest at uniform pressure p0 pushes each straight side of the unit
   square with p0 times its outward normal, and the closed contour with zero.
 * Plane Couette flow u = U y (uniform p and T, so the entropy-variable field is
   linear in the velocity components and the vertex gradients are exact): the
   viscous force on the wall at rest is mu U per unit length along x, on the
   moving wall -mu U; the pressure force is p0 times the outward normal.
"""
import numpy as np

import tensum_amd as T
import oracle_binding as O
from test_oracle_walls import couette_mesh


def test_uniform_pressure_on_the_sides_and_the_closed_contour():
    m = T.jittered_mesh(17, seed=4, bc={t: ("slip",) for t in (1, 2, 3, 4)})
    o = O.Oracle(m, T.Params(flux="kepec", reconstruct="first", cfl=0.5, final_time=1.0))
    p0 = 1.7
    o.set_state(np.tile([1.2, 0.0, 0.0, p0], (m.n_vertex, 1)))
    for tag, n in ((1, (0, -1)), (2, (1, 0)), (3, (0, 1)), (4, (-1, 0))):
        f = o.forces([tag])
        np.testing.assert_allclose(f[:2], p0 * np.array(n), atol=1e-14)
        assert np.all(f[2:] == 0.0)
    assert np.abs(o.forces([1, 2, 3, 4])[:2]).max() < 1e-14


def test_couette_shear_on_the_walls():
    U, mu = 0.4, 0.05
    m = couette_mesh(17, U, 1.0)
    # adiabatic walls: the wall states keep the fluid's uniform temperature
    m.bc[1] = ("noslip", (1.0, 0.0, 0.0, 1.0), True)
    m.bc[3] = ("noslip", (1.0, U, 0.0, 1.0), True)
    o = O.Oracle(m, T.Params(model="ns", flux="kepec", reconstruct="first", mu_ref=mu,
                             cfl=0.5, final_time=1.0))
    o.set_state(np.stack([np.ones_like(m.y), U * m.y, 0 * m.y, np.ones_like(m.y)], 1))
    bottom = int(np.flatnonzero(m.bface_tag == 1)[0])
    top = int(np.flatnonzero(m.bface_tag == 3)[0])
    fb = o.forces([1], force_index=bottom)
    ft = o.forces([3], force_index=top)
    np.testing.assert_allclose(fb, [0.0, -1.0, mu * U, 0.0], atol=1e-12)
    np.testing.assert_allclose(ft, [0.0, 1.0, -mu * U, 0.0], atol=1e-12)
