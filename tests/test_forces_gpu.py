"""Force integrals (src/force.cc:118-217) through tsm_forces against the oracle,
Euler and Navier-Stokes, on the initial state and after steps of the tiled and
CSR kernels."""
import numpy as np
import pytest

import tensum_amd as T
import oracle_binding as O
from cases import rel_l2

pytestmark = pytest.mark.gpu

ST = (1.0, 0.5, 0.0, 1.0)


def _bump(m):
    g = np.exp(-40.0 * ((m.x - 0.45) ** 2 + (m.y - 0.5) ** 2))
    return np.stack([1.0 + 0.1 * g, 0.5 + 0.05 * g, 0.03 * g, 1.0 + 0.2 * g], 1)


CASES = {
    "euler_slip": (dict(flux="kepes_tecno_roe", reconstruct="minmod"),
                   {1: ("slip",), 3: ("slip",), 4: ("inlet", ST), 2: ("outlet",)}),
    "ns_walls": (dict(model="ns", flux="kepes_tecno_roe", reconstruct="minmod", mu_ref=0.01),
                 {1: ("noslip", (1.1, 0.0, 0.0, 1.0)), 3: ("noslip", (1.0, 0.0, 0.0, 1.0), True),
                  4: ("farfield", ST), 2: ("outlet",)}),
}


@pytest.mark.parametrize("path", ["auto", "csr"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_forces_match_oracle(case, path):
    kw, bc = CASES[case]
    m = T.jittered_mesh(33, seed=12, bc=bc)
    prm = T.Params(cfl=0.5, final_time=10.0, kernel_path=path, **kw)
    s, o = T.Solver(m, prm), O.Oracle(m, prm)
    s.set_state(_bump(m))
    o.set_state(_bump(m))
    wall = int(np.flatnonzero(m.bface_tag == 1)[2])
    for steps in (0, 15, 10):
        s.run(steps)
        o.run(steps)
        for tags, fi in (([1, 3], 0), ([1], wall), ([4], 5), ([1, 2, 3, 4], wall)):
            a, b = s.forces(tags, force_index=fi), o.forces(tags, force_index=fi)
            np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-13 * np.abs(b).max())


def test_bad_force_requests_are_refused():
    m = T.jittered_mesh(9, seed=12, bc=CASES["euler_slip"][1])
    s = T.Solver(m, T.Params(flux="kepec", reconstruct="first", cfl=0.5, final_time=1.0))
    s.set_state(_bump(m))
    with pytest.raises(T.TsmError):
        s.forces([99])
    with pytest.raises(T.TsmError):
        s.forces([1], force_index=10 ** 6)


@pytest.mark.parametrize("path", ["auto", "csr"])
def test_single_sample_queries_of_a_batch(path):
    """Per-sample queries after batched steps (each forms only its own
    sample's entropy variables and gradients): gradients, Navier-Stokes
    forces and current-state surfaces of every sample match its own oracle."""
    kw, bc = CASES["ns_walls"]
    m = T.jittered_mesh(25, seed=14, bc=bc)
    prm = T.Params(cfl=0.5, final_time=10.0, kernel_path=path, **kw)
    amps = (0.5, 1.0, 1.5)
    s = T.Solver(m, prm, n_samples=len(amps))
    for k, a in enumerate(amps):
        s.set_state(_bump(m) * np.array([1.0, a, 1.0, 1.0]), k)
    s.run(12)
    wall = int(np.flatnonzero(m.bface_tag == 1)[2])
    for k in (2, 0, 1):
        o = O.Oracle(m, prm)
        o.set_state(_bump(m) * np.array([1.0, amps[k], 1.0, 1.0]))
        o.run(12)
        assert rel_l2(s.gradients(k), o.gradients()) <= 1e-10
        np.testing.assert_allclose(s.forces([1, 3], sample=k, force_index=wall),
                                   o.forces([1, 3], force_index=wall), rtol=1e-10, atol=1e-14)
        assert rel_l2(s.surface([1, 3], sample=k, stage=False), o.surface([1, 3])) <= 1e-10
