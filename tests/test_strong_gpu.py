"""bc_scheme strong on the GPU (CSR kernels) against the oracle: the boundary
values imposed after every stage in bface order (src/fv.cc:1694-1704), the
entropy variables left as the update made them, the last stage's momentum
residual zeroed on noslip end points (src/fv.cc:1805-1819).  Euler and
Navier-Stokes walls, the three Runge-Kutta schemes, steady mode, per-sample
boundary values and a periodic channel.  Per-step logs to 1e-12, states to
rel-L2 1e-10 in conserved variables.
"""
import numpy as np
import pytest

import tensum_amd as T
import oracle_binding as O
from cases import conserved, rel_l2
from test_oracle_bc_values import smooth_values
from test_oracle_walls import couette_mesh

pytestmark = pytest.mark.gpu

ST = (1.0, 0.4, 0.05, 1.0)
BC = {1: ("noslip", (1.3, 0.0, 0.0, 1.0)), 3: ("noslip", (1.0, 0.1, 0.0, 1.0), True),
      4: ("inlet", ST), 2: ("pressure", ST)}
BC2 = {1: ("slip",), 3: ("farfield", ST), 4: ("inlet", ST), 2: ("outlet",)}


def _state(m, amp=0.1):
    return np.stack([1.0 + amp * np.sin(np.pi * m.x), 0.4 + amp * m.y, 0.05 * np.cos(np.pi * m.x),
                     1.0 + 0.5 * amp * np.sin(np.pi * m.y)], 1)


def _prm(**kw):
    return T.Params(**{**dict(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.5, final_time=10.0,
                              bc_scheme="strong"), **kw})


CASES = {
    "euler_first": (BC, _prm(flux="kepec", reconstruct="first")),
    "euler_minmod": (BC, _prm()),
    "euler_open": (BC2, _prm(reconstruct="van_albada")),
    "euler_rk4": (BC, _prm(time_scheme="rk4")),
    "euler_heun": (BC, _prm(time_scheme="heuns")),
    "ns_walls": (BC, _prm(model="ns", mu_ref=0.01)),
    "ns_rk4_sutherland": (BC, _prm(model="ns", mu_ref=0.02, viscosity="sutherland", T_0=0.4,
                                   time_scheme="rk4")),
    "ns_steady": (BC, _prm(model="ns", mu_ref=0.01, time_mode="steady", max_iter=10 ** 6)),
    "euler_voronoi": (BC2, _prm(cell="voronoi")),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_strong_runs_match_oracle(case):
    bc, prm = CASES[case]
    m = T.jittered_mesh(25, seed=8, bc=bc)
    s, o = T.Solver(m, prm), O.Oracle(m, prm)
    s.set_state(_state(m))
    o.set_state(_state(m))
    done, lg = s.run(30, want_log=True)
    n, lo = o.run(30, want_log=True)
    assert done[0] == n == 30
    d = np.abs(lg[:, 1:] - lo[:, 1:]) / np.abs(lo[:, 1:])
    assert d.max() <= 1e-12, d.max(axis=0)
    assert rel_l2(conserved(s.get_state()), conserved(o.get_state())) <= 1e-10


def test_strong_batched_samples_with_their_own_boundary_values():
    m = T.jittered_mesh(21, seed=9, bc=BC)
    prm = _prm(model="ns", mu_ref=0.01)
    s = T.Solver(m, prm, n_samples=3)
    vals = [smooth_values(m, a) for a in (0.0, 0.1, 0.2)]
    for k, v in enumerate(vals):
        s.set_bc_state(v, k)
        s.set_state(_state(m, 0.05 * (k + 1)), k)
    s.run(20)
    for k, v in enumerate(vals):
        o = O.Oracle(m, prm)
        o.set_bc_state(v)
        o.set_state(_state(m, 0.05 * (k + 1)))
        o.run(20)
        assert rel_l2(conserved(s.get_state(k)), conserved(o.get_state())) <= 1e-10


def test_strong_periodic_couette_channel():
    m = couette_mesh(17, 0.4, 1.0)
    prm = _prm(model="ns", flux="kepec", reconstruct="first", mu_ref=0.05, cfl=0.8)
    w = np.stack([1.0 + 0 * m.y, 0.4 * m.y, 0 * m.y, 1.0 + 0.01 * np.sin(2 * np.pi * m.x)], 1)
    s, o = T.Solver(m, prm), O.Oracle(m, prm)
    s.set_state(w)
    o.set_state(w)
    s.run(40)
    o.run(40)
    assert rel_l2(conserved(s.get_state()), conserved(o.get_state())) <= 1e-10
    W = s.get_state()
    wall = np.unique(np.asarray(m.bface_v).reshape(-1, 2)[np.asarray(m.bface_tag) == 1].ravel())
    assert np.all(W[wall, 1] == 0.0) and np.all(W[wall, 0] == 1.0)
