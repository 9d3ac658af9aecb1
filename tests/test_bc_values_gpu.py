"""Per-sample, per-end-point boundary values (tsm_set_bc_state) on the GPU
against the oracle with the same values: the Monte Carlo path of the
reference's UQ examples, whose boundary expressions carry each sample's
random numbers (src/bc.h:411-562, src/perturb.cc).  Batched samples with
different values each match their own oracle run; tiled and CSR kernels;
Euler and Navier-Stokes walls; graph replay after the values change.
"""
import numpy as np
import pytest

import tensum_amd as T
import oracle_binding as O
from cases import conserved, rel_l2
from test_oracle_bc_values import BC, _state, smooth_values, table_values

pytestmark = pytest.mark.gpu


def _prm(**kw):
    return T.Params(**{**dict(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.5,
                              final_time=10.0), **kw})


CASES = {
    "euler_tiled": _prm(),
    "euler_csr": _prm(kernel_path="csr"),
    "euler_fused": _prm(fused_gradient=True),
    "ns_walls": _prm(model="ns", mu_ref=0.01),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_batched_samples_with_their_own_boundary_values(case):
    prm = CASES[case]
    m = T.jittered_mesh(25, seed=2, bc=BC)
    vals = [smooth_values(m, amp) for amp in (0.05, 0.1, 0.15)]
    s = T.Solver(m, prm, n_samples=3)
    for k, v in enumerate(vals):
        s.set_bc_state(v, k)
        s.set_state(_state(m), k)
    R = [s.residual(k) for k in range(3)]
    s.run(25)
    for k, v in enumerate(vals):
        o = O.Oracle(m, prm)
        o.set_bc_state(v)
        o.set_state(_state(m))
        assert rel_l2(R[k], o.residual()) < 1e-12
        o.run(25)
        assert rel_l2(conserved(s.get_state(k)), conserved(o.get_state())) <= 1e-10


@pytest.mark.parametrize("path", ["auto", "csr"])
def test_table_values_leave_the_run_unchanged(path):
    m = T.jittered_mesh(21, seed=3, bc=BC)
    out = []
    for vals in (None, table_values(m)):
        s = T.Solver(m, _prm(kernel_path=path), n_samples=2)
        for k in range(2):
            if vals is not None:
                s.set_bc_state(vals, k)
            s.set_state(_state(m), k)
        s.run(15)
        out.append([s.get_state(k) for k in range(2)])
    for k in range(2):
        assert np.array_equal(out[0][k], out[1][k])


def test_graph_replay_sees_new_boundary_values():
    """A graph captured before the samples got their own values is recaptured;
    replays then equal host-launched steps."""
    m = T.jittered_mesh(25, seed=4, bc=BC)
    g = T.Solver(m, _prm(graph_steps=4), n_samples=2)
    e = T.Solver(m, _prm(), n_samples=2)
    for s in (g, e):
        for k in range(2):
            s.set_state(_state(m), k)
        s.run(8)
        s.set_bc_state(smooth_values(m, 0.12), 1)
        s.run(8)
        s.set_bc_state(smooth_values(m, 0.02), 0)
        s.run(8)
    for k in range(2):
        assert np.array_equal(g.get_state(k), e.get_state(k))


def test_bad_boundary_values_are_refused():
    m = T.jittered_mesh(13, seed=2, bc=BC)
    s = T.Solver(m, _prm())
    ib = int(np.flatnonzero(m.bface_tag == 4)[0])
    for bad in (0.0, -1.0, np.nan):
        v = table_values(m)
        v[ib, 1, 0] = bad
        with pytest.raises(T.TsmError):
            s.set_bc_state(v)
    with pytest.raises(T.TsmError):
        s.set_bc_state(table_values(m)[:-1])
    with pytest.raises(T.TsmError):
        s.set_bc_state(table_values(m), 1)
