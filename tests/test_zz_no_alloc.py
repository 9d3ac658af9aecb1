"""tsm_probe_density_noalloc (the fixed-scratch variant for callers that cannot
allocate device memory) against tsm_probe_density, bit for bit: two calls in a
row on one context, the second with fewer probes (the scratch array keeps the
first call's contents), and a second context of another mesh; the size limit
and points outside the mesh are refused before or without a wrong result."""
import numpy as np
import pytest

import tensum_amd as T
from cases import vortex, vortex_mesh, riemann, riemann_mesh

pytestmark = pytest.mark.gpu


def _solver(mesh, states, prm):
    s = T.Solver(mesh, prm, n_samples=len(states))
    for k, w in enumerate(states):
        s.set_state(w, k)
    return s


def test_noalloc_probe_density_equals_the_allocating_entry_point():
    print("noalloc: vortex 33^2, 3 samples", flush=True)
    m = vortex_mesh(33)
    prm = T.Params(flux="kepes_tecno_roe", reconstruct="minmod", cfl=0.8, final_time=10.0)
    s = _solver(m, [vortex(m.x, m.y, beta=b) for b in (4.0, 5.0, 6.0)], prm)
    s.run(5)
    rng = np.random.default_rng(7)
    px, py = rng.uniform(-4.9, 4.9, 64), rng.uniform(-4.9, 4.9, 64)
    a = s.probe_density(px, py)
    b = s.probe_density_noalloc(px, py)
    assert np.all(np.isfinite(a))
    assert np.array_equal(a, b)
    # second call, fewer probes: nothing of the first call's contents leaks in
    a2 = s.probe_density(px[:9][::-1], py[:9][::-1])
    b2 = s.probe_density_noalloc(px[:9][::-1], py[:9][::-1])
    assert b2.shape == (3, 9) and np.array_equal(a2, b2)
    # mesh vertices themselves (points on cell edges and corners)
    a3 = s.probe_density(m.x[::97], m.y[::97])
    b3 = s.probe_density_noalloc(m.x[::97], m.y[::97])
    assert np.array_equal(a3, b3)


def test_noalloc_probe_density_second_context_and_errors():
    print("noalloc: riemann 65^2, limits", flush=True)
    m = riemann_mesh(65)
    prm = T.Params(flux="kepes_tecno_rusanov", reconstruct="minmod", cfl=0.5, final_time=0.8)
    s = _solver(m, [riemann(m.x, m.y)], prm)
    s.run(3)
    px = np.linspace(0.01, 0.99, 50)
    py = np.linspace(0.99, 0.01, 50)
    assert np.array_equal(s.probe_density(px, py), s.probe_density_noalloc(px, py))
    with pytest.raises(T.TsmError, match="outside the mesh"):
        s.probe_density_noalloc(np.array([0.5, 1.5]), np.array([0.5, 0.5]))
    # after a refused call the next one is still right
    assert np.array_equal(s.probe_density(px[:5], py[:5]), s.probe_density_noalloc(px[:5], py[:5]))
    big = np.full(400000, 0.5)  # 400,000 probes x (1 sample) need more than 8 MiB
    with pytest.raises(T.TsmError, match="scratch"):
        s.probe_density_noalloc(big, big)
