"""Boundary values of time on the GPU (tsm_set_bc_time_expr): the device
evaluates each boundary's expressions of (x, y, t) at every end point before
each Heun stage at the sample's stage time t^n + r dt, which the reference
passes to BoundaryCondition::apply (src/fv.cc:2296-2298, src/bc.h:411-562).
Compared with the oracle, whose callback gives it the same fields from numpy:
tiled, CSR and fused-gradient kernels, Navier-Stokes walls, batched samples at
different times, graph replay.  The device's sin/exp/tanh and the host's may
differ in the last bit, so states are compared to the 1e-10 relative-L2 bar.
"""
import numpy as np
import pytest

import tensum_amd as T
import oracle_binding as O
from cases import conserved, rel_l2
from test_oracle_bc_time import BC, EXPR, callback, prm, state

pytestmark = pytest.mark.gpu

TAGS = sorted(EXPR)


def _groups(m):
    return np.array([TAGS.index(int(t)) for t in m.bface_tag], dtype=np.int32)


def _exprs():
    return [EXPR[t] for t in TAGS]


CASES = {
    "tiled": prm(),
    "csr": prm(kernel_path="csr"),
    "fused": prm(fused_gradient=True),
    "ns_walls": prm(model="ns", mu_ref=0.01),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_boundary_values_of_time_match_the_oracle(case):
    p = CASES[case]
    m = T.jittered_mesh(25, seed=11, bc=BC)
    W = [state(m), state(m)[::-1].copy()]
    W[1][:, 1] *= -1.0
    clocks = [(0, 0, 0.0, 0.0), (40, 1, 0.25, 1.0)]
    s = T.Solver(m, p, n_samples=2)
    for k in range(2):
        s.set_state(W[k], k)
        if clocks[k][0]:
            s.set_clock(*clocks[k], sample=k)
    s.set_bc_time_expr(_groups(m), _exprs())
    s.run(30)
    for k in range(2):
        o = O.Oracle(m, p)
        o.set_state(W[k])
        if clocks[k][0]:
            o.set_clock(*clocks[k])
        fn, calls = callback(m)
        o.set_bc_time_fn(fn)
        assert o.run(30) == 30
        assert len(calls) == 60
        got = s.get_state(k)
        assert np.isfinite(got).all()
        assert rel_l2(conserved(got), conserved(o.get_state())) <= 1e-10, (case, k)
        it, _, t, _ = s.clock(k)
        assert it == o.clock()[0]
        assert t == pytest.approx(o.clock()[2], rel=1e-13)


def test_expressions_without_t_equal_set_bc_state():
    """Fields that do not depend on t, given as expressions, run as the same
    values given per end point (arithmetic only, so numpy forms them bit for
    bit)."""
    m = T.jittered_mesh(21, seed=12, bc=BC)
    ex = {4: ("1.0 + 0.1*x*y", "0.4 - 0.05*y", None, "1.0 + 0.02*(x + y)"),
          3: (None, "0.4 + 0.1*x", None, None), 2: (None, None, None, "0.98 + 0.01*x"),
          1: ("1.2 - 0.1*x", None, None, None)}
    vals = np.zeros((m.bface_tag.size, 2, 4))
    for b, tag in enumerate(m.bface_tag):
        vals[b, :, :] = m.bc[int(tag)][1]
    bv = m.bface_v.reshape(-1, 2)
    for e in range(2):
        x, y = m.x[bv[:, e]], m.y[bv[:, e]]
        f = {4: (1.0 + 0.1 * x * y, 0.4 - 0.05 * y, None, 1.0 + 0.02 * (x + y)),
             3: (None, 0.4 + 0.1 * x, None, None), 2: (None, None, None, 0.98 + 0.01 * x),
             1: (1.2 - 0.1 * x, None, None, None)}
        for tag in ex:
            sel = m.bface_tag == tag
            for q in range(4):
                if f[tag][q] is not None:
                    vals[sel, e, q] = f[tag][q][sel]
    out = []
    for use_expr in (False, True):
        s = T.Solver(m, prm(), n_samples=2)
        for k in range(2):
            s.set_state(state(m), k)
            if not use_expr:
                s.set_bc_state(vals, k)
        if use_expr:
            s.set_bc_time_expr(_groups(m), [ex[t] for t in TAGS])
        s.run(12)
        out.append([s.get_state(k) for k in range(2)])
    for k in range(2):
        assert np.array_equal(out[0][k], out[1][k])


def test_graph_replay_equals_launched_steps():
    m = T.jittered_mesh(25, seed=13, bc=BC)
    res = []
    for gs in (0, 4):
        s = T.Solver(m, prm(graph_steps=gs), n_samples=2)
        for k in range(2):
            s.set_state(state(m), k)
        s.set_bc_time_expr(_groups(m), _exprs())
        s.run(16)
        res.append([s.get_state(k) for k in range(2)])
    for k in range(2):
        assert np.array_equal(res[0][k], res[1][k])


def test_removing_the_expressions_restores_constant_values():
    m = T.jittered_mesh(17, seed=14, bc=BC)
    a = T.Solver(m, prm())
    a.set_state(state(m))
    a.set_bc_time_expr(_groups(m), [("1 + 0*t", None, None, None)] * len(TAGS))
    a.set_bc_time_expr(_groups(m), [])
    a.run(5)
    # the values the expressions last wrote stay until replaced; here they
    # equal the table's T except on the wall (1.2)
    b = T.Solver(m, prm())
    b.set_state(state(m))
    vals = np.zeros((m.bface_tag.size, 2, 4))
    for i, tag in enumerate(m.bface_tag):
        vals[i, :, :] = m.bc[int(tag)][1]
        vals[i, :, 0] = 1.0
    b.set_bc_state(vals)
    b.run(5)
    assert np.array_equal(a.get_state(), b.get_state())


def test_refusals():
    m = T.jittered_mesh(9, seed=15, bc=BC)
    g = _groups(m)
    with pytest.raises(T.TsmError, match="heuns"):
        T.Solver(m, prm(time_scheme="ssprk3")).set_bc_time_expr(g, _exprs())
    s = T.Solver(m, prm())
    with pytest.raises(T.TsmError, match="pert"):
        s.set_bc_time_expr(g, [("1 + pert(1, x, y, 0)*t", None, None, None)] * len(TAGS))
    with pytest.raises(T.TsmError, match="unknown function"):
        s.set_bc_time_expr(g, [("1 + foo(t)", None, None, None)] * len(TAGS))
    with pytest.raises(T.TsmError, match="arguments"):
        s.set_bc_time_expr(g, [("1 + pow(t)", None, None, None)] * len(TAGS))
    with pytest.raises(T.TsmError, match="group"):
        s.set_bc_time_expr(g + 7, _exprs())
