"""Median-dual geometry built on the device (face discovery by sorting edge
keys, the reference's face numbering and orientation, normals, cell and dual
areas, periodic sums; src/grid_preproc.cc:18-590) against the host builder,
bit for bit, and solver runs on either preprocessing."""
import numpy as np
import pytest

import tensum_amd as T
from cases import riemann, riemann_mesh, tgv_mesh, vortex, vortex_mesh
from test_gpu_parity import _fan_mesh
from test_oracle_walls import couette_mesh

pytestmark = pytest.mark.gpu


def _shuffled(m, seed=0):
    """Same mesh with cells in another order, each cell's vertices rotated,
    and every other boundary face listed backwards."""
    rng = np.random.default_rng(seed)
    cells = m.cell_v.reshape(-1, 3)[rng.permutation(m.cell_v.size // 3)]
    rot = rng.integers(0, 3, cells.shape[0])
    cells = np.stack([np.roll(c, -r) for c, r in zip(cells, rot)])
    bv = m.bface_v.reshape(-1, 2).copy()
    bv[::2] = bv[::2, ::-1]
    return T.Mesh(m.x, m.y, m.bface_tag, bv.ravel(), cells.ravel().astype(np.int32),
                  m.periodic_ptr, m.periodic_v, m.bc)


MESHES = {
    "riemann": lambda: riemann_mesh(65),
    "vortex_periodic": lambda: vortex_mesh(65),
    "fan": lambda: _fan_mesh(12),
    "couette": lambda: couette_mesh(33, 0.4, 1.0),
    "shuffled": lambda: _shuffled(T.jittered_mesh(41, seed=3, bc={1: ("slip",), 2: ("outlet",),
                                                             3: ("slip",), 4: ("outlet",)})),
    "unjittered": lambda: T.jittered_mesh(17, jitter=0.0),
    # the benchmark's mesh: 1,050,625 vertices, 2,097,152 cells
    "riemann_1025": lambda: riemann_mesh(1025),
    # the periodic Taylor-Green mesh of the Navier-Stokes configuration
    "tgv_513": lambda: tgv_mesh(513),
}


@pytest.mark.parametrize("kind", sorted(MESHES))
def test_device_geometry_is_the_host_geometry(kind):
    m = MESHES[kind]()
    h, d = T.host_geometry(m), T.device_geometry(m)
    for k in h:
        assert np.array_equal(h[k], d[k]), k


@pytest.mark.parametrize("kind", sorted(MESHES))
def test_device_tiling_is_the_host_tiling(kind):
    """Tiles (periodic sets first, then Morton order), halos (ascending),
    local faces and boundary faces, gather and cell entries, periodic slot
    tables, numbering: built on the device by key sorts and binary searches,
    equal to the host's build entry for entry."""
    m = MESHES[kind]()
    for tv in ((64, 128, 256) if m.n_vertex < 10 ** 5 else (128,)):
        n, which = T.device_tiling_check(m, tv)
        assert n == 0, (kind, tv, which)


def test_edges_the_host_repairs_are_left_to_it():
    """An outer edge no boundary face names is a one-cell interior face, which
    the host builder drops; the device builder hands such meshes back."""
    m = riemann_mesh(17)
    keep = np.ones(m.bface_tag.size, bool)
    keep[5] = False
    cut = T.Mesh(m.x, m.y, m.bface_tag[keep], m.bface_v.reshape(-1, 2)[keep].ravel(), m.cell_v,
                 m.periodic_ptr, m.periodic_v, m.bc)
    with pytest.raises(T.TsmError):
        T.device_geometry(cut)
    prm = T.Params(flux="kepes_tecno_rusanov", reconstruct="minmod", cfl=0.5, final_time=0.8)
    out = []
    for pre in ("auto", "host"):
        s = T.Solver(cut, T.Params(**{**prm.__dict__, "preprocess": pre}))
        s.set_state(riemann(cut.x, cut.y))
        s.run(5)
        out.append(s.get_state())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("path", ["auto", "csr"])
@pytest.mark.parametrize("periodic", [False, True])
def test_runs_on_device_and_host_preprocessing_are_identical(path, periodic):
    m = vortex_mesh(49) if periodic else riemann_mesh(49)
    w = vortex(m.x, m.y) if periodic else riemann(m.x, m.y)
    out = []
    for pre in ("auto", "host"):
        prm = T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.5, final_time=10.0,
                       kernel_path=path, preprocess=pre)
        s = T.Solver(m, prm)
        s.set_state(w)
        s.run(20)
        out.append((s.get_state(), s.geometry()["dcarea"]))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
