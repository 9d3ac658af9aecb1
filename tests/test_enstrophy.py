"""Enstrophy, 1/2 sum omega^2 dcarea w, with the reference's vorticity from
entropy-variable gradients (Material::Vorticity, src/material.h:622-627).

The reference integrates kinetic energy and entropy (compute_global,
src/fv.cc:2175-2225) but no enstrophy; the Navier-Stokes Taylor-Green
workload reports it.  Pins of the oracle's restatement: the Taylor-Green
vortex u = sin x cos y, v = -cos x sin y has omega = 2 sin x sin y, so
1/2 int omega^2 = 2 pi^2 on [0, 2pi]^2, approached at second order in h; a
rigid rotation u = -W y, v = W x at constant T and p has linear entropy
variables, whose Green-Gauss gradients are exact, so omega = 2W at every
vertex and the enstrophy is 2 W^2 times the domain area to rounding.
"""
import numpy as np
import pytest

import oracle_binding as O
import tensum_amd as T
from cases import rel_l2, riemann_mesh, taylor_green, tgv_mesh

PRM = T.Params(model="ns", flux="kepec", reconstruct="first", mu_ref=1.0 / 1600, prandtl=0.72,
               cfl=0.5, final_time=10.0)


def _rotation(mesh, W=0.7):
    w = np.empty((mesh.n_vertex, 4))
    w[:, 0] = 1.3
    w[:, 1] = -W * (mesh.y - 0.5)
    w[:, 2] = W * (mesh.x - 0.5)
    w[:, 3] = 2.0
    return w


def test_oracle_enstrophy_taylor_green_converges_to_2pi2():
    errs = []
    for n in (17, 33, 65):
        m = tgv_mesh(n)
        o = O.Oracle(m, PRM)
        o.set_state(taylor_green(m.x, m.y))
        errs.append(abs(o.enstrophy() / (2 * np.pi ** 2) - 1))
    assert errs[-1] < 1e-2
    for a, b in zip(errs, errs[1:]):
        assert 3.5 < a / b < 4.5  # second order


def test_oracle_enstrophy_of_a_rigid_rotation_is_exact():
    m = riemann_mesh(21)
    o = O.Oracle(m, T.Params())
    o.set_state(_rotation(m))
    assert abs(o.enstrophy() - 2 * 0.7 ** 2) <= 1e-13


@pytest.mark.gpu
def test_gpu_enstrophy_matches_oracle():
    m = tgv_mesh(65)
    s = T.Solver(m, PRM)
    o = O.Oracle(m, PRM)
    s.set_state(taylor_green(m.x, m.y))
    o.set_state(taylor_green(m.x, m.y))
    assert abs(s.enstrophy() - o.enstrophy()) <= 1e-12 * o.enstrophy()
    s.run(15)
    o.run(15)
    assert abs(s.enstrophy() - o.enstrophy()) <= 1e-11 * o.enstrophy()
    r = riemann_mesh(21)
    s = T.Solver(r, T.Params())
    s.set_state(_rotation(r))
    assert abs(s.enstrophy() - 2 * 0.7 ** 2) <= 1e-13


@pytest.mark.gpu
def test_driver_writes_enstrophy_column(tmp_path):
    """tensum -e: global.dat gains a fifth column, the enstrophy after each
    step (the reference's four columns unchanged)."""
    import os
    import subprocess
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    from cases import write_part_file
    from ref_cases import CASES, param_text
    c = CASES["tgv33_ns"]
    mesh = c["mesh"]()
    write_part_file(mesh, str(tmp_path), periodic_tags=(1, 2, 3, 4))
    (tmp_path / "param.in").write_text(param_text("tgv33_ns", 0.05, 10 ** 6))
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tensum_amd",
                       "tensum")
    r = subprocess.run([exe, "-i", "param.in", "-npart", "1", "-L", ".", "-e"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = np.loadtxt(tmp_path / "SAMPLE_0" / "global.dat", ndmin=2)
    assert g.shape[1] == 5
    o = O.Oracle(mesh, T.Params(final_time=0.05, **c["params"]))
    o.set_state(c["state"](mesh))
    n = o.run(10 ** 5)
    assert g.shape[0] == n
    np.testing.assert_allclose(g[-1, 2:], [*o.global_integrals(), o.enstrophy()], rtol=6e-6)
