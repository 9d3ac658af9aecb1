"""Pins of the oracle's Navier-Stokes wall cells (src/fv.cc:1236-1312,
src/material.cc:236-302) and of adiabatic noslip walls (src/bc.h:226-276,
:411-451).

 * The wall flux of a non-adiabatic wall is the cell flux through the wall
   normal, bit for bit; an adiabatic wall's is the same flux with heat
   conduction switched off (the correction term cancels exactly the
   conduction share), and at a wall at rest it carries no energy at all.
 * A closed box with adiabatic noslip walls at rest conserves mass and total
   energy (no mass or energy crosses a wall at rest that conducts no heat)
   while friction takes kinetic energy out; a fluid at rest stays at rest.
 * Isothermal walls hotter than the fluid heat it; the box stays at rest when
   the wall temperature equals the fluid's.
 * Plane Couette flow between a wall at rest and a moving isothermal wall
   (periodic in x): the steady state is u = U y/H, v = 0, constant p, and the
   quadratic temperature of viscous heating; the steady discrete solution
   converges to it at second order.
"""
import numpy as np
import pytest

import tensum_amd as T
import oracle_binding as O
from cases import conserved

GAM, R = 1.4, 1.0


def _ns(**kw):
    base = dict(model="ns", flux="kepec", reconstruct="first", mu_ref=0.02, prandtl=0.72,
                cfl=0.5, final_time=100.0)
    return T.Params(**{**base, **kw})


def _random_state(rng):
    T_, u, v = 1.0 + 0.3 * rng.random(), rng.standard_normal(), rng.standard_normal()
    p = 1.0 + 0.5 * rng.random()
    rho = p / (R * T_)
    s = np.log(p) - GAM * np.log(rho)
    beta = 1.0 / (R * T_)
    ent = [(GAM - s) / (GAM - 1) - beta * (u * u + v * v) / 2, beta * u, beta * v, -beta]
    return np.array(ent), rng.standard_normal(8), rng.standard_normal(2)


def test_non_adiabatic_wall_flux_is_the_cell_flux():
    rng = np.random.default_rng(0)
    for _ in range(20):
        ent, dE, n = _random_state(rng)
        a = O.viscous_flux(_ns(), ent, dE, n, wall=True, adiabatic=False)
        b = O.viscous_flux(_ns(), ent, dE, n)
        assert np.array_equal(a, b)


def test_adiabatic_wall_flux_has_no_heat_conduction():
    rng = np.random.default_rng(1)
    for _ in range(20):
        ent, dE, n = _random_state(rng)
        a = O.viscous_flux(_ns(), ent, dE, n, wall=True, adiabatic=True)
        cell = O.viscous_flux(_ns(), ent, dE, n)
        cold = O.viscous_flux(_ns(heat_active=False), ent, dE, n)
        assert np.array_equal(a[:3], cell[:3])
        assert abs(a[3] - cold[3]) <= 1e-13 * (abs(cell[3]) + abs(cold[3]))
        assert abs(cell[3] - cold[3]) > 1e-6  # conduction is really there in the cell flux
    # a wall at rest that conducts no heat passes no energy
    ent, dE, n = _random_state(rng)
    ent[1:3] = 0.0
    a = O.viscous_flux(_ns(), ent, dE, n, wall=True, adiabatic=True)
    cell = O.viscous_flux(_ns(), ent, dE, n)
    assert abs(a[3]) <= 1e-14 * abs(cell[3]) and a[0] == 0.0


def _box(n, wall):
    return T.jittered_mesh(n, box=(0.0, 1.0, 0.0, 1.0), jitter=0.2, seed=5,
                           bc={t: wall for t in (1, 2, 3, 4)})


def _totals(o):
    a = o.geometry()["dcarea"]
    q = conserved(o.get_state())
    return (q * a[:, None]).sum(0)


def test_closed_adiabatic_box_conserves_mass_and_energy():
    m = _box(17, ("noslip", (1.0, 0.0, 0.0, 1.0), True))
    x, y = m.x, m.y
    w = np.stack([1.0 + 0.2 * np.sin(np.pi * x) * np.sin(np.pi * y),
                  0.3 * np.sin(np.pi * x) * np.sin(2 * np.pi * y),
                  -0.3 * np.sin(2 * np.pi * x) * np.sin(np.pi * y),
                  1.0 + 0.1 * np.cos(np.pi * x)], 1)
    o = O.Oracle(m, _ns())
    o.set_state(w)
    q0 = _totals(o)
    ke0 = (0.5 * conserved(w)[:, 0] * (w[:, 1] ** 2 + w[:, 2] ** 2) * o.geometry()["dcarea"]).sum()
    assert o.run(150) == 150
    q1 = _totals(o)
    np.testing.assert_allclose(q1[[0, 3]], q0[[0, 3]], rtol=1e-13)
    W = o.get_state()
    ke1 = (0.5 * conserved(W)[:, 0] * (W[:, 1] ** 2 + W[:, 2] ** 2) * o.geometry()["dcarea"]).sum()
    assert ke1 < 0.9 * ke0


@pytest.mark.parametrize("Tw", [1.0, 1.3])
def test_isothermal_walls_heat_a_fluid_at_rest(Tw):
    m = _box(13, ("noslip", (Tw, 0.0, 0.0, 1.0)))
    w = np.tile([1.0, 0.0, 0.0, 1.0], (m.n_vertex, 1))
    o = O.Oracle(m, _ns())
    o.set_state(w)
    q0 = _totals(o)
    o.run(60)
    q1 = _totals(o)
    np.testing.assert_allclose(q1[0], q0[0], rtol=1e-13)
    if Tw == 1.0:
        assert np.abs(o.get_state() - w).max() <= 1e-13
    else:
        assert q1[3] > q0[3] * (1 + 1e-4)  # heat flows in through the walls
    # the same box with adiabatic walls leaves the fluid at rest
    a = O.Oracle(_box(13, ("noslip", (Tw, 0.0, 0.0, 1.0), True)), _ns())
    a.set_state(w)
    a.run(60)
    assert np.abs(a.get_state() - w).max() <= 1e-13


def couette_mesh(n, U, Tw, seed=9):
    """[0,1]^2, periodic in x (left/right joined), wall at rest at y = 0 and a
    wall moving with U at y = 1, both isothermal at Tw."""
    m = T.jittered_mesh(n, box=(0.0, 1.0, 0.0, 1.0), jitter=0.2, seed=seed,
                        bc={1: ("noslip", (Tw, 0.0, 0.0, 1.0)), 3: ("noslip", (Tw, U, 0.0, 1.0)),
                            2: ("periodic",), 4: ("periodic",)})
    sets = [[j * n, j * n + n - 1] for j in range(n)]   # left and right column, same y
    ptr = np.cumsum([0] + [len(s) for s in sets]).astype(np.int32)
    return T.Mesh(m.x, m.y, m.bface_tag, m.bface_v, m.cell_v, ptr,
                  np.array(sum(sets, []), dtype=np.int32), m.bc)


def couette_exact(y, U, Tw, mass, pr=0.72):
    """Steady plane Couette flow in the unit channel: u = U y, v = 0,
    k T'' + mu U^2 = 0 with T(0) = T(1) = Tw and k = mu cp / pr, and the
    uniform pressure that holds the channel's mass (p int 1/(R T) = mass)."""
    cp = GAM * R / (GAM - 1)
    a = pr * U * U / (2 * cp)
    Tm = Tw + a * y * (1 - y)
    # int_0^1 dy / (Tw + a y (1 - y)), Simpson on a fine grid
    t = np.linspace(0.0, 1.0, 2001)
    f = 1.0 / (R * (Tw + a * t * (1 - t)))
    integral = (f[0] + f[-1] + 4 * f[1:-1:2].sum() + 2 * f[2:-1:2].sum()) / (3 * 2000)
    return np.stack([Tm, U * y, 0 * y, mass / integral + 0 * y], 1)


@pytest.mark.slow_cpu
def test_couette_flow_converges_at_second_order():
    U, Tw, mu = 0.4, 1.0, 0.05
    errs = []
    for n in (9, 17):
        m = couette_mesh(n, U, Tw)
        prm = _ns(mu_ref=mu, cfl=0.8, time_mode="steady", min_residue=1e-9, max_iter=40000)
        o = O.Oracle(m, prm)
        o.set_state(np.tile([Tw, 0.0, 0.0, 1.0], (m.n_vertex, 1)))  # mass 1 in the channel
        o.run(40000)
        assert o.run(1) == 0 and o.time() > 0  # converged to min_residue, not cut off
        W = o.get_state()
        ex = couette_exact(m.y, U, Tw, 1.0)
        errs.append(np.sqrt(np.mean((W - ex) ** 2, 0)))
    errs = np.array(errs)
    rate = np.log2(errs[0] / errs[1])
    # T, u, v, p: second order in u and T; v and the pressure level are far smaller
    assert np.all(errs[1] < [1e-4, 3e-5, 3e-6, 1e-4]), errs
    assert rate[0] > 1.7 and rate[1] > 1.7, (errs, rate)
