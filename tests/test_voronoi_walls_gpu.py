"""Navier-Stokes wall cells on Voronoi duals from which preprocessing removed
empty faces (src/grid_preproc.cc:633-647, the reference's only Navier-Stokes
example uses Voronoi cells): the cells' face slots keep their indices into
the shortened list (src/grid_preproc.cc:811-831); boundary faces come first
and are never removed, so the wall cells read their own boundary edges.  Runs,
forces and surfaces against the oracle; tiled and CSR kernels; strong and
weak boundary conditions."""
import numpy as np
import pytest

import tensum_amd as T
import oracle_binding as O
from cases import conserved, rel_l2

pytestmark = pytest.mark.gpu

ST = (1.0, 0.4, 0.05, 1.0)
BC = {1: ("noslip", (1.3, 0.0, 0.0, 1.0)), 3: ("noslip", (1.0, 0.1, 0.0, 1.0), True),
      4: ("inlet", ST), 2: ("pressure", ST)}


def _state(m):
    return np.stack([1.0 + 0.1 * np.sin(np.pi * m.x), 0.4 + 0.1 * m.y, 0.05 * np.cos(np.pi * m.x),
                     1.0 + 0.05 * np.sin(np.pi * m.y)], 1)


@pytest.mark.parametrize("path,scheme", [("auto", "weak"), ("csr", "weak"), ("auto", "strong")])
def test_voronoi_wall_cells_match_oracle(path, scheme):
    m = T.jittered_mesh(25, seed=8, bc=BC)
    nf = {c: T.host_geometry(m, c)["face_v"].shape[0] for c in ("median", "voronoi")}
    assert nf["voronoi"] < nf["median"]   # preprocessing removed empty faces
    prm = T.Params(model="ns", mu_ref=0.01, flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.5,
                   final_time=10.0, cell="voronoi", kernel_path=path, bc_scheme=scheme)
    s, o = T.Solver(m, prm), O.Oracle(m, prm)
    s.set_state(_state(m))
    o.set_state(_state(m))
    assert rel_l2(s.residual(), o.residual()) < 1e-12
    done, lg = s.run(30, want_log=True)
    n, lo = o.run(30, want_log=True)
    assert done[0] == n == 30
    assert (np.abs(lg[:, 1:] - lo[:, 1:]) / np.abs(lo[:, 1:])).max() <= 1e-12
    assert rel_l2(conserved(s.get_state()), conserved(o.get_state())) <= 1e-10
    wall = int(np.flatnonzero(m.bface_tag == 1)[1])
    np.testing.assert_allclose(s.forces([1, 3], force_index=wall), o.forces([1, 3], force_index=wall),
                               rtol=1e-10, atol=1e-14)
    assert rel_l2(s.surface([1, 3], stage=False), o.surface([1, 3])) <= 1e-10


def test_unjittered_lattice_with_every_hypotenuse_removed():
    """An unjittered lattice: the hypotenuse faces between right triangles are
    empty Voronoi faces, all removed; closed box of isothermal walls."""
    m = T.jittered_mesh(9, jitter=0.0, bc={t: ("noslip", (1.0, 0.0, 0.0, 1.0)) for t in (1, 2, 3, 4)})
    nf = {c: T.host_geometry(m, c)["face_v"].shape[0] for c in ("median", "voronoi")}
    assert nf["voronoi"] < nf["median"]
    prm = T.Params(model="ns", flux="kepec", reconstruct="first", mu_ref=0.01, cfl=0.5,
                   final_time=10.0, cell="voronoi")
    w = np.stack([1.0 + 0.1 * m.x, 0.05 * np.sin(np.pi * m.y), 0 * m.x, 1.0 + 0.02 * m.y], 1)
    s, o = T.Solver(m, prm), O.Oracle(m, prm)
    s.set_state(w)
    o.set_state(w)
    s.run(20)
    o.run(20)
    assert rel_l2(conserved(s.get_state()), conserved(o.get_state())) <= 1e-10
