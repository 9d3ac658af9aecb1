"""Pins of the oracle's strong boundary conditions (bc_scheme strong):

 * One Heun step on a first-order Euler run (no gradients, so the entropy
   variables the update leaves behind play no part) equals the procedure of
   src/fv.cc:1694-1704 and :1805-1819 written out here from the oracle's
   residual: each stage's update, then every bface in bface order imposing
   its boundary condition on its two end points (noslip: wall velocity and
   temperature; slip: the normal velocity removed; inlet: the whole state;
   pressure: p; farfield and outlet: nothing); in the last stage the
   momentum residual of noslip end points is zeroed before the update.
 * After any step the imposed values hold exactly at vertices whose bfaces
   all carry one boundary: wall velocity and temperature on isothermal noslip
   walls, the inlet state at inlet vertices.
 * The weak scheme is the default and differs from the strong one.
"""
import numpy as np

import tensum_amd as T
import oracle_binding as O
from cases import conserved, rel_l2
from test_oracle_heun import _prim

ST = (1.0, 0.4, 0.05, 1.0)
WALL = (1.3, 0.0, 0.0, 1.0)
BC = {1: ("noslip", WALL), 3: ("slip",), 4: ("inlet", ST), 2: ("pressure", ST)}


def _state(m):
    return np.stack([1.0 + 0.1 * np.sin(np.pi * m.x), 0.4 + 0.1 * m.y, 0.05 * np.cos(np.pi * m.x),
                     1.0 + 0.05 * np.sin(np.pi * m.y)], 1)


def _impose(W, geo, tags):
    W = W.copy()
    for (v0, v1), n, t in zip(geo["bface_v"], geo["bface_n"], tags):
        for v in (v0, v1):
            kind = BC[int(t)][0]
            if kind == "noslip":
                W[v, 0], W[v, 1], W[v, 2] = WALL[0], WALL[1], WALL[2]
            elif kind == "slip":
                u = n / np.sqrt(n @ n)
                f = W[v, 1] * u[0] + W[v, 2] * u[1]
                W[v, 1] -= u[0] * f
                W[v, 2] -= u[1] * f
            elif kind == "inlet":
                W[v] = ST
            elif kind == "pressure":
                W[v, 3] = ST[3]
    return W


def test_strong_heun_step_is_the_reference_procedure():
    m = T.jittered_mesh(17, seed=3, bc=BC)
    prm = T.Params(flux="kepec", reconstruct="first", cfl=0.5, final_time=10.0, time_scheme="heuns",
                   bc_scheme="strong")
    o = O.Oracle(m, prm)
    geo = o.geometry()
    tags = np.asarray(m.bface_tag)
    A = geo["dcarea"][:, None]
    W0 = _state(m)
    o.set_state(W0)
    dt = o.compute_dt()
    U0 = conserved(W0)
    W1 = _impose(_prim(U0 - (dt / A) * o.residual()), geo, tags)
    o.set_state(W1)
    R1 = o.residual()
    noslip = np.unique(geo["bface_v"][tags == 1].ravel())
    R1[noslip, 1:3] = 0.0
    W2 = _impose(_prim(0.5 * U0 + 0.5 * (conserved(W1) - (dt / A) * R1)), geo, tags)
    s = O.Oracle(m, prm)
    s.set_state(W0)
    assert s.run(1) == 1
    assert rel_l2(conserved(s.get_state()), conserved(W2)) < 1e-13
    # the momentum zeroing matters: without it the step differs
    R1b = o.residual()
    W2b = _impose(_prim(0.5 * U0 + 0.5 * (conserved(W1) - (dt / A) * R1b)), geo, tags)
    assert rel_l2(conserved(W2b), conserved(W2)) > 1e-8


def test_imposed_values_hold_after_steps():
    m = T.jittered_mesh(17, seed=4, bc=BC)
    for extra in (dict(), dict(model="ns", mu_ref=0.01, flux="kepes_tecno_roe", reconstruct="minmod"),
                  dict(time_scheme="rk4")):
        prm = T.Params(**{**dict(flux="kepec", reconstruct="first", cfl=0.5, final_time=10.0,
                                 bc_scheme="strong"), **extra})
        o = O.Oracle(m, prm)
        o.set_state(_state(m))
        o.run(12)
        W = o.get_state()
        geo = o.geometry()
        tags = np.asarray(m.bface_tag)
        bv = geo["bface_v"]
        only = lambda t: np.setdiff1d(np.unique(bv[tags == t].ravel()), np.unique(bv[tags != t].ravel()))
        assert np.all(W[only(1)][:, :3] == np.array(WALL[:3]))
        assert np.all(W[only(4)] == np.array(ST))
        assert np.all(W[only(2)][:, 3] == ST[3])
        weak = O.Oracle(m, T.Params(**{**prm.__dict__, "bc_scheme": "weak"}))
        weak.set_state(_state(m))
        weak.run(12)
        assert rel_l2(conserved(weak.get_state()), conserved(W)) > 1e-6
