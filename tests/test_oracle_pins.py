"""Further pins of the oracle (CPU), each fixed by mathematics or by data the
reference itself holds:

* free-stream preservation on open domains whose inlet, farfield, pressure,
  outlet and slip sides hold the free-stream state (src/bc.h:569-654,
  src/material.cc:110-127, src/fv.cc:1100-1189): a closed dual cell sums its
  face normals to zero, so R = 0 for every flux;
* known answers of the Liou carbuncle sensor (src/compute_ssw.cc:7-39) on an
  unjittered lattice (every dual-face normal then points from the face's
  left to its right vertex): a stationary normal shock (supersonic into
  subsonic) raises it exactly at the end points of the faces the shock
  crosses, an expansion and a state at rest raise it nowhere (on jittered
  meshes a dual normal can point against its edge, and the sensor then reads
  an expansion along it as a compression, as the reference's test is written);
* closed forms of the global integrals of a uniform state
  (src/fv.cc:2175-2225): KE = 1/2 rho |u|^2 A, S = -rho s/(gamma-1) A with
  A the domain area and s = ln p - gamma ln rho;
* closed forms of two fluxes: KEP's energy flux is rho_bar H_bar vn_bar with
  arithmetic means and H = c^2/(gamma-1) + |u|^2/2 (src/kep.cc:10-38), and
  simple_avg is the mean of the two exact Euler fluxes (src/simple_avg.cc);
* the reference's own exact shock-tube profile
  (examples/Euler/shock_tube/ref_soln.dat, copied to tests/golden/
  sod_ref_soln.dat: x, rho, u, p, e at t = 0.2): the example's physics (left
  state T 1, u 0.75, p 1; right T 0.8, p 0.1; split at x = 0.3; inlet and
  outlet ends, periodic sides) on strip meshes converges to it in L1, and the
  profile itself is the exact Riemann solution (checked against an
  independent exact solver here).
"""
import os

import numpy as np
import pytest

import oracle_binding as O
import tensum_amd as T
from cases import riemann_mesh, vortex_mesh

HERE = os.path.dirname(os.path.abspath(__file__))
FLUXES = ["kepes_tecno_roe", "kepes_tecno_rusanov", "kepec", "roe_ec", "kep", "simple_avg"]


# ---- free stream on open domains -------------------------------------------
@pytest.mark.parametrize("flux", FLUXES)
@pytest.mark.parametrize("sides", ["slip_outlet_inlet", "farfield_pressure_inlet"])
def test_free_stream_preserved_with_open_boundaries(flux, sides):
    if sides == "slip_outlet_inlet":  # flow along the slip walls
        W = (1.3, 0.7, 0.0, 1.1)
        bc = {1: ("slip",), 2: ("outlet",), 3: ("slip",), 4: ("inlet", W)}
    else:  # oblique flow through farfield sides, pressure outlet
        W = (1.3, 0.7, -0.4, 1.1)
        bc = {1: ("farfield", W), 2: ("pressure", W), 3: ("farfield", W), 4: ("inlet", W)}
    m = T.jittered_mesh(19, seed=7, bc=bc)
    o = O.Oracle(m, T.Params(flux=flux, reconstruct="minmod", cfl=0.5, final_time=10.0))
    w = np.tile(np.array(W), (m.n_vertex, 1))
    o.set_state(w)
    assert np.abs(o.residual()).max() < 1e-13
    o.run(20)
    np.testing.assert_allclose(o.get_state(), w, rtol=1e-13, atol=1e-13)


# ---- Liou sensor ---------------------------------------------------------------
def _shock_state(mesh, left, right, x0=0.525):
    w = np.empty((mesh.n_vertex, 4))
    w[mesh.x < x0] = left
    w[mesh.x >= x0] = right
    return w


def _crossing(o, mesh, x0=0.525):
    g = o.geometry()
    fv, fn = g["face_v"], g["face_n"]
    xl, xr = mesh.x[fv[:, 0]], mesh.x[fv[:, 1]]
    cross = (xl < x0) != (xr < x0)
    return fv, fn, cross


def _lattice(n=21):
    return T.jittered_mesh(n, jitter=0.0, seed=1, bc={t: ("outlet",) for t in (1, 2, 3, 4)})


def test_liou_sensor_flags_a_normal_shock_only_across_it():
    m = _lattice()
    o = O.Oracle(m, T.Params(flux="kepes_tecno_roe", liou_fix=True))
    # Mach 2 flow along +x into Mach 0.5 flow (T = 1, c = sqrt(1.4)): both sides'
    # u.n - c|n| change sign across the shock for normals along x
    c = np.sqrt(1.4)
    o.set_state(_shock_state(m, (1.0, 2.0 * c, 0.0, 1.0), (1.0, 0.5 * c, 0.0, 2.0)))
    ssw = o.ssw()
    fv, fn, cross = _crossing(o, m)
    along_x = np.abs(fn[:, 0]) > 0.9 * np.hypot(fn[:, 0], fn[:, 1])
    must = np.unique(fv[cross & along_x].ravel())
    may = np.unique(fv[cross].ravel())
    assert must.size > 0 and np.all(ssw[must] == 1)
    assert set(np.flatnonzero(ssw)) <= set(may)
    g = o.geometry()  # the lattice's dual normals point from left to right vertex
    d = np.stack([m.x[fv[:, 1]] - m.x[fv[:, 0]], m.y[fv[:, 1]] - m.y[fv[:, 0]]], 1)
    assert np.all(np.sum(d * g["face_n"], 1) > 0)
    # the same shock moving the other way (flow along -x), raised through u.n + c|n|
    o.set_state(_shock_state(m, (1.0, -0.5 * c, 0.0, 2.0), (1.0, -2.0 * c, 0.0, 1.0)))
    ssw = o.ssw()
    assert np.all(ssw[must] == 1) and set(np.flatnonzero(ssw)) <= set(may)


def test_liou_sensor_stays_off_for_expansions_and_rest():
    m = _lattice()
    o = O.Oracle(m, T.Params(flux="kepes_tecno_roe", liou_fix=True))
    c = np.sqrt(1.4)
    o.set_state(_shock_state(m, (1.0, 0.5 * c, 0.0, 2.0), (1.0, 2.0 * c, 0.0, 1.0)))  # expansion
    assert o.ssw().sum() == 0
    o.set_state(_shock_state(m, (1.0, 0.0, 0.0, 2.0), (1.0, 0.0, 0.0, 1.0)))  # at rest
    assert o.ssw().sum() == 0


# ---- integrals of a uniform state ----------------------------------------------
@pytest.mark.parametrize("periodic", [True, False])
def test_global_integrals_of_a_uniform_state(periodic):
    m = vortex_mesh(17) if periodic else riemann_mesh(17)
    area = 100.0 if periodic else 1.0
    Tm, u, v, p, gam = 1.3, 0.7, -0.4, 1.1, 1.4
    o = O.Oracle(m, T.Params())
    o.set_state(np.tile([Tm, u, v, p], (m.n_vertex, 1)))
    rho = p / Tm
    s = np.log(p) - gam * np.log(rho)
    ke, ent = o.global_integrals()
    assert ke == pytest.approx(0.5 * rho * (u * u + v * v) * area, rel=1e-13)
    assert ent == pytest.approx(-rho * s / (gam - 1) * area, rel=1e-13)


# ---- KEP and simple_avg closed forms ---------------------------------------------
def _euler(W, n, gam=1.4):
    Tm, u, v, p = W
    rho = p / Tm
    vn = u * n[0] + v * n[1]
    E = p / (gam - 1) + 0.5 * rho * (u * u + v * v)
    return np.array([rho * vn, rho * u * vn + p * n[0], rho * v * vn + p * n[1], (E + p) * vn])


PAIRS = [((1.0, 0.3, -0.2, 1.0), (1.3, 0.1, 0.4, 0.7)),
         ((0.2, 1.206, 1.206, 0.029), (0.56, 0.0, 1.206, 0.3))]


def _flux(name, Wl, Wr, n):
    z = np.zeros(8)
    return O.num_flux(T.Params(flux=name, reconstruct="first"), np.zeros(2), np.ones(2),
                      np.array(Wl, float), np.array(Wr, float), z, z, np.array(n, float))


@pytest.mark.parametrize("Wl,Wr", PAIRS)
def test_kep_energy_flux_closed_form(Wl, Wr):
    n = np.array([0.6, 0.8])
    F = _flux("kep", Wl, Wr, n)
    gam = 1.4

    def H(W):
        Tm, u, v, p = W
        return gam * Tm / (gam - 1) + 0.5 * (u * u + v * v)  # c^2 = gamma R T, R = 1
    rho = 0.5 * (Wl[3] / Wl[0] + Wr[3] / Wr[0])
    vn = 0.5 * (np.array(Wl[1:3]) + np.array(Wr[1:3])) @ n
    assert F[3] == pytest.approx(rho * 0.5 * (H(Wl) + H(Wr)) * vn, rel=1e-14)


@pytest.mark.parametrize("Wl,Wr", PAIRS)
def test_simple_avg_is_the_mean_of_the_euler_fluxes(Wl, Wr):
    n = np.array([0.6, 0.8])
    F = _flux("simple_avg", Wl, Wr, n)
    np.testing.assert_allclose(F, 0.5 * (_euler(Wl, n) + _euler(Wr, n)), rtol=1e-14)


# ---- the shock-tube example against the reference's exact profile ----------------
def _sod_exact(x, t, WL, WR, x0, gam=1.4):
    """Exact Riemann solution (two-rarefaction/shock Newton iteration on p*),
    primitive (rho, u, p) at x, t."""
    rl, ul, pl = WL
    rr, ur, pr = WR
    cl, cr = np.sqrt(gam * pl / rl), np.sqrt(gam * pr / rr)

    def f(p, r, pk, ck):
        if p > pk:
            A, B = 2 / ((gam + 1) * r), (gam - 1) / (gam + 1) * pk
            return (p - pk) * np.sqrt(A / (p + B))
        return 2 * ck / (gam - 1) * ((p / pk) ** ((gam - 1) / (2 * gam)) - 1)
    lo, hi = 1e-8, 10.0
    for _ in range(200):  # bisection on the pressure function
        mid = 0.5 * (lo + hi)
        if f(mid, rl, pl, cl) + f(mid, rr, pr, cr) + ur - ul > 0:
            hi = mid
        else:
            lo = mid
    ps = 0.5 * (lo + hi)
    us = 0.5 * (ul + ur) + 0.5 * (f(ps, rr, pr, cr) - f(ps, rl, pl, cl))
    out = np.empty((x.size, 3))
    for i, s in enumerate((x - x0) / t):
        if s < us:  # left of the contact
            if ps > pl:
                rsl = rl * (ps / pl + (gam - 1) / (gam + 1)) / ((gam - 1) / (gam + 1) * ps / pl + 1)
                sl = ul - cl * np.sqrt((gam + 1) / (2 * gam) * ps / pl + (gam - 1) / (2 * gam))
                out[i] = (rl, ul, pl) if s < sl else (rsl, us, ps)
            else:
                rsl = rl * (ps / pl) ** (1 / gam)
                csl = cl * (ps / pl) ** ((gam - 1) / (2 * gam))
                if s < ul - cl:
                    out[i] = (rl, ul, pl)
                elif s > us - csl:
                    out[i] = (rsl, us, ps)
                else:
                    c = 2 / (gam + 1) * (cl + (gam - 1) / 2 * (ul - s))
                    u = 2 / (gam + 1) * (cl + (gam - 1) / 2 * ul + s)
                    r = rl * (c / cl) ** (2 / (gam - 1))
                    out[i] = (r, u, pl * (c / cl) ** (2 * gam / (gam - 1)))
        else:  # right of the contact (shock or rarefaction)
            if ps > pr:
                rsr = rr * (ps / pr + (gam - 1) / (gam + 1)) / ((gam - 1) / (gam + 1) * ps / pr + 1)
                sr = ur + cr * np.sqrt((gam + 1) / (2 * gam) * ps / pr + (gam - 1) / (2 * gam))
                out[i] = (rr, ur, pr) if s > sr else (rsr, us, ps)
            else:
                rsr = rr * (ps / pr) ** (1 / gam)
                csr = cr * (ps / pr) ** ((gam - 1) / (2 * gam))
                if s > ur + cr:
                    out[i] = (rr, ur, pr)
                elif s < us + csr:
                    out[i] = (rsr, us, ps)
                else:
                    c = 2 / (gam + 1) * (cr - (gam - 1) / 2 * (ur - s))
                    u = 2 / (gam + 1) * (-cr + (gam - 1) / 2 * ur + s)
                    r = rr * (c / cr) ** (2 / (gam - 1))
                    out[i] = (r, u, pr * (c / cr) ** (2 * gam / (gam - 1)))
    return out


def _ref_profile():
    return np.loadtxt(os.path.join(HERE, "golden", "sod_ref_soln.dat"))


def test_reference_profile_is_the_exact_riemann_solution():
    ref = _ref_profile()
    ex = _sod_exact(ref[:, 0], 0.2, (1.0, 0.75, 1.0), (0.125, 0.0, 0.1), 0.3)
    # printed at 6 decimals; the cells that straddle a jump may differ
    ok = np.abs(ref[:, 1] - ex[:, 0]) < 2e-6
    assert ok.mean() > 0.99
    np.testing.assert_allclose(ref[ok, 1:4], ex[ok], atol=2e-6)


def _strip(nx, ny=4):
    """nx x ny vertex strip on [0,1] x [0, ny-1)/(nx-1)], squares split along one
    diagonal, counter-clockwise triangles; tags 4 left (inlet), 2 right
    (outlet), 1 bottom and 3 top joined periodically."""
    h = 1.0 / (nx - 1)
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    x, y = (ix * h).ravel(), (iy * h).ravel()
    vid = lambda i, j: j * nx + i
    cells = []
    for j in range(ny - 1):
        for i in range(nx - 1):
            a, b, c, d = vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)
            cells += [(a, b, c), (a, c, d)]
    bt, bv = [], []
    for i in range(nx - 1):
        bt += [1, 3]
        bv += [(vid(i, 0), vid(i + 1, 0)), (vid(i + 1, ny - 1), vid(i, ny - 1))]
    for j in range(ny - 1):
        bt += [2, 4]
        bv += [(vid(nx - 1, j), vid(nx - 1, j + 1)), (vid(0, j + 1), vid(0, j))]
    pp, pv = [0], []
    for i in range(nx):
        pv += [vid(i, 0), vid(i, ny - 1)]
        pp.append(len(pv))
    Wl = (1.0, 0.75, 0.0, 1.0)
    bc = {1: ("periodic",), 3: ("periodic",), 2: ("outlet",), 4: ("inlet", Wl)}
    return T.Mesh(x, y, np.array(bt, np.int32), np.array(bv, np.int32).ravel(),
                  np.array(cells, np.int32).ravel(), np.array(pp, np.int32),
                  np.array(pv, np.int32), bc)


def test_shock_tube_converges_to_the_reference_profile():
    ref = _ref_profile()
    errs = []
    for nx in (51, 101, 201):
        m = _strip(nx)
        o = O.Oracle(m, T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.5,
                                 final_time=0.2))
        w = np.empty((m.n_vertex, 4))
        left = m.x < 0.3
        w[left] = (1.0, 0.75, 0.0, 1.0)
        w[~left] = (0.8, 0.0, 0.0, 0.1)
        o.set_state(w)
        o.run(10 ** 6)
        assert o.time() == pytest.approx(0.2, abs=1e-14)
        st = o.get_state()
        row = m.y == 0.0
        xs, rho = m.x[row], st[row, 3] / st[row, 0]
        errs.append(np.mean(np.abs(np.interp(ref[:, 0], xs, rho) - ref[:, 1])))
    assert errs[-1] < 0.02
    assert errs[0] > errs[1] > errs[2]
    assert errs[0] / errs[2] > 2.5  # about first order with the shock and the contact
