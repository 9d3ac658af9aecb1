"""Navier-Stokes wall cells (src/fv.cc:1236-1312) and adiabatic noslip walls
(src/bc.h:226-276, :411-451) on the GPU against the oracle.  Wall cells run
on the CSR kernels (the tiled ones take meshes without them); the Euler cases
with adiabatic walls run on the tiled kernels.  Tolerances as elsewhere:
per-call residuals to rel-L2 1e-12, runs to 1e-10 in conserved variables,
logs to 1e-12.
"""
import numpy as np
import pytest

import tensum_amd as T
import oracle_binding as O
from cases import conserved, rel_l2
from test_oracle_walls import couette_mesh

pytestmark = pytest.mark.gpu

ST = (1.0, 0.3, 0.05, 1.0)


def _ns(**kw):
    base = dict(model="ns", flux="kepes_tecno_roe", reconstruct="minmod", mu_ref=0.01,
                prandtl=0.72, cfl=0.5, final_time=100.0)
    return T.Params(**{**base, **kw})


def _wavy(m):
    x, y = m.x, m.y
    return np.stack([1.0 + 0.1 * np.sin(np.pi * x) * np.cos(np.pi * y),
                     0.3 + 0.1 * np.sin(2 * np.pi * y), 0.05 * np.cos(np.pi * x),
                     1.0 + 0.05 * np.cos(np.pi * (x + y))], 1)


def _run_pair(m, prm, w, steps):
    s, o = T.Solver(m, prm), O.Oracle(m, prm)
    s.set_state(w)
    o.set_state(w)
    assert rel_l2(s.residual(), o.residual()) < 1e-12
    done, lg = s.run(steps, want_log=True)
    n, lo = o.run(steps, want_log=True)
    assert done[0] == n == steps
    d = np.abs(lg[:, 1:] - lo[:, 1:]) / np.abs(lo[:, 1:])
    assert d.max() <= 1e-12, d.max(axis=0)
    assert rel_l2(conserved(s.get_state()), conserved(o.get_state())) <= 1e-10
    return s, o


WALLS = {
    "adiabatic_box": {t: ("noslip", (1.0, 0.0, 0.0, 1.0), True) for t in (1, 2, 3, 4)},
    "isothermal_box": {t: ("noslip", (1.2, 0.0, 0.0, 1.0)) for t in (1, 2, 3, 4)},
    "channel": {1: ("noslip", (1.1, 0.0, 0.0, 1.0)), 3: ("slip",), 4: ("inlet", ST),
                2: ("outlet",)},
    "farfield_pressure": {1: ("noslip", (1.0, 0.0, 0.0, 1.0), True), 2: ("pressure", ST),
                          3: ("farfield", ST), 4: ("farfield", ST)},
}


@pytest.mark.parametrize("walls", sorted(WALLS))
@pytest.mark.parametrize("flux,rec", [("kepes_tecno_roe", "minmod"), ("kepec", "first")])
def test_navier_stokes_wall_cells_match_oracle(walls, flux, rec):
    m = T.jittered_mesh(25, seed=7, bc=WALLS[walls])
    _run_pair(m, _ns(flux=flux, reconstruct=rec), _wavy(m), 40)


def test_navier_stokes_wall_cells_sutherland_no_heat():
    m = T.jittered_mesh(21, seed=8, bc=WALLS["channel"])
    _run_pair(m, _ns(viscosity="sutherland", mu_ref=0.02, T_0=0.4, heat_active=False), _wavy(m),
              30)


def test_steady_couette_matches_oracle():
    m = couette_mesh(17, 0.4, 1.0)
    prm = _ns(flux="kepec", reconstruct="first", mu_ref=0.05, cfl=0.8, time_mode="steady",
              max_iter=10 ** 6)
    w = np.tile([1.0, 0.0, 0.0, 1.0], (m.n_vertex, 1))
    s, o = _run_pair(m, prm, w, 300)
    k1, e1 = s.global_integrals()
    k2, e2 = o.global_integrals()
    assert abs(k1 - k2) <= 1e-10 * abs(k2)


@pytest.mark.parametrize("path", ["auto", "csr"])
def test_euler_adiabatic_noslip_matches_oracle(path):
    """Adiabatic noslip walls keep the fluid temperature in both boundary
    applies; on the tiled kernels (auto) and the CSR kernels."""
    bc = {1: ("noslip", (5.0, 0.0, 0.0, 1.0), True), 2: ("outlet",), 3: ("slip",),
          4: ("inlet", ST)}
    m = T.jittered_mesh(29, seed=4, bc=bc)
    prm = T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.5, final_time=10.0,
                   kernel_path=path)
    s, o = _run_pair(m, prm, _wavy(m), 40)
    # the wall's temperature value plays no part when it is adiabatic
    bc2 = dict(bc)
    bc2[1] = ("noslip", (0.7, 0.0, 0.0, 1.0), True)
    s2 = T.Solver(T.jittered_mesh(29, seed=4, bc=bc2), prm)
    s2.set_state(_wavy(m))
    s2.run(40)
    assert np.array_equal(s2.get_state(), s.get_state())
