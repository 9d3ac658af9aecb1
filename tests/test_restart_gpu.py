"""Restart continuation on the GPU (tsm_set_clock, src/fv.cc:300-402): a run
stopped after k steps and continued in a new context from its state and loop
clock is the uninterrupted run bit for bit; tiled, graph-replayed and CSR
steps, unsteady with time stamps and steady; each sample of a batch with its
own clock; bad clocks refused."""
import numpy as np
import pytest

import tensum_amd as T
import oracle_binding as O
from cases import vortex, vortex_mesh

pytestmark = pytest.mark.gpu

CASES = {
    "tiled": T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=0.6,
                      time_stamps=4),
    "graph": T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=0.6,
                      time_stamps=4, graph_steps=3),
    "csr_rk4": T.Params(flux="kepec", reconstruct="van_albada", cfl=0.8, final_time=0.6,
                        time_stamps=3, kernel_path="csr", time_scheme="rk4"),
    "steady": T.Params(flux="kepec", reconstruct="first", cfl=0.8, time_mode="steady", max_iter=40,
                       min_residue=1e-12),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_continuation_is_the_uninterrupted_run(case):
    prm = CASES[case]
    m = vortex_mesh(17)
    a = T.Solver(m, prm)
    a.set_state(vortex(m.x, m.y))
    n = int(a.run(100000)[0])
    o = O.Oracle(m, prm)
    o.set_state(vortex(m.x, m.y))
    assert o.run(100000) == n
    assert np.abs(a.get_state() - o.get_state()).max() <= 1e-10 * np.abs(o.get_state()).max()
    for k in (3, n - 2):
        b = T.Solver(m, prm)
        b.set_state(vortex(m.x, m.y))
        assert b.run(k)[0] == k
        clock = b.clock()
        c = T.Solver(m, prm)
        c.set_state(b.get_state())
        c.set_clock(*clock)
        assert c.run(100000)[0] == n - k
        assert np.array_equal(c.get_state(), a.get_state())
        assert c.clock() == a.clock()


def test_batched_samples_with_their_own_clocks():
    prm = CASES["tiled"]
    m = vortex_mesh(17)
    full = T.Solver(m, prm)
    full.set_state(vortex(m.x, m.y))
    n = int(full.run(100000)[0])
    ks = (2, 5, 9)
    c = T.Solver(m, prm, n_samples=len(ks))
    for j, k in enumerate(ks):
        b = T.Solver(m, prm)
        b.set_state(vortex(m.x, m.y))
        b.run(k)
        c.set_state(b.get_state(), j)
        c.set_clock(*b.clock(), sample=j)
    done = c.run(100000)
    for j, k in enumerate(ks):
        assert done[j] == n - k
        assert np.array_equal(c.get_state(j), full.get_state())


def test_bad_clocks_are_refused():
    m = vortex_mesh(9)
    s = T.Solver(m, CASES["tiled"])
    s.set_state(vortex(m.x, m.y))
    for bad in ((-1, 1, 0.1, 1.0), (3, 99, 0.1, 1.0), (3, 1, np.nan, 1.0), (3, 1, 0.1, 0.0),
                (3, 1, -0.5, 1.0)):
        with pytest.raises(T.TsmError):
            s.set_clock(*bad)
    s.set_clock(3, 2, 0.6, 1.0)      # at the final time: the run is over
    assert s.run(5)[0] == 0


def test_bounds_are_the_state_extremes():
    """tsm_bounds (src/fv.cc:2099-2171) equals numpy's extremes of the state,
    exactly, for each sample of a batch and a mesh of more vertices than one
    pass of its reduction blocks covers."""
    m = vortex_mesh(601)
    s = T.Solver(m, CASES["tiled"], n_samples=2)
    for k in range(2):
        s.set_state(vortex(m.x, m.y) * (1.0 + 0.1 * k), k)
    s.run(3)
    for k in range(2):
        W = s.get_state(k)
        want = [W[:, 0].min(), W[:, 0].max(), W[:, 3].min(), W[:, 3].max(), W[:, 1].min(),
                W[:, 1].max(), W[:, 2].min(), W[:, 2].max(), 0.0, 0.0]
        assert np.array_equal(s.bounds(k), np.array(want))


def test_status_all_equals_per_sample_status():
    m = vortex_mesh(17)
    s = T.Solver(m, CASES["tiled"], n_samples=3)
    for k in range(3):
        s.set_state(vortex(m.x, m.y) * (1.0 + 0.2 * k), k)
    s.run(7)
    a = s.status_all()
    for k in range(3):
        one = s.status(k)
        for key in ("t", "dt", "res", "iter", "counter", "done"):
            assert a[key][k] == one[key]
        assert np.array_equal(a["norms"][k], s.residual_norms(k))
