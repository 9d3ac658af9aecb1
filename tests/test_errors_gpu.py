"""Error paths of the C ABI: bad settings, bad meshes and bad states come back
as TSM_E_* codes with a message (the reference aborts or exits instead), and
finished samples stay finished.  Every case fails on the host before any
kernel sees the bad input."""
import numpy as np
import pytest

import tensum_amd as T
from cases import vortex, vortex_mesh

pytestmark = pytest.mark.gpu

BASE = dict(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=1.0)


def _mesh():
    return vortex_mesh(9)


@pytest.mark.parametrize("change,needle", [
    ({"gamma": 1.0}, "gamma"),
    ({"time_mode": "steady", "max_iter": 0}, "max_iter"),
    ({"min_residue": -1.0}, "min_residue"),
    ({"cfl": 0.0}, "time_step or cfl"),
    ({"time_step": 0.1}, "both time_step and cfl"),
    ({"final_time": 0.0}, "final_time"),
    ({"time_stamps": 0}, "time_stamps"),
    ({"mu_ref": 0.1}, "mu_ref"),
    ({"tile_vertices": 300}, "tile_vertices"),
    ({"graph_steps": -1}, "graph_steps"),
])
def test_bad_settings_are_refused(change, needle):
    with pytest.raises(T.TsmError, match=needle):
        T.Solver(_mesh(), T.Params(**{**BASE, **change}))


def test_bad_meshes_are_refused():
    m = _mesh()
    cells = m.cell_v.copy()
    cells[4] = m.n_vertex + 3
    with pytest.raises(T.TsmError, match="out of range"):
        T.Solver(T.Mesh(**{**m.__dict__, "cell_v": cells}), T.Params(**BASE))
    flat = m.cell_v.reshape(-1, 3).copy()
    flat[0] = flat[0][[0, 0, 1]]  # degenerate cell
    with pytest.raises(T.TsmError):
        T.Solver(T.Mesh(**{**m.__dict__, "cell_v": flat.ravel()}), T.Params(**BASE))
    with pytest.raises(T.TsmError):
        T.Solver(m, T.Params(**BASE), n_samples=0)


def test_bad_states_and_indices_are_refused():
    m = _mesh()
    s = T.Solver(m, T.Params(**BASE), n_samples=2)
    w = vortex(m.x, m.y)
    s.set_state(w, 0)
    s.set_state(w, 1)
    for col, bad in ((0, np.nan), (0, -1.0), (3, 0.0), (3, np.inf), (1, np.nan), (2, -np.inf)):
        v = w.copy()
        v[3, col] = bad
        with pytest.raises(T.TsmError):
            s.set_state(v, 1)
    with pytest.raises(T.TsmError):
        s.set_state(w, 2)
    with pytest.raises(T.TsmError):
        s.get_state(-1)
    with pytest.raises(T.TsmError):
        s.set_material(0, 1.0, 1.0)
    assert np.array_equal(s.get_state(1), w)  # a refused call leaves the state alone


def test_finished_samples_stay_finished():
    m = _mesh()
    s = T.Solver(m, T.Params(**{**BASE, "final_time": 0.05}))
    s.set_state(vortex(m.x, m.y))
    assert s.run(0)[0] == 0
    n = int(s.run(10000)[0])
    assert n > 0 and s.status()["done"]
    w = s.get_state()
    assert int(s.run(50)[0]) == 0
    dt, _ = s.step()
    assert dt[0] == 0.0
    assert np.array_equal(s.get_state(), w)
