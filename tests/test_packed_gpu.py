"""Packed multilevel steps (tsm_run_packed): one launch per kernel spanning
the tiles and samples of several contexts with meshes of different sizes
must give each context exactly what its own launches give (same per-block
work, same summation order), for periodic meshes (in-tile periodic sums) and
for meshes with boundary-corrected gradients (the boundary-vertex pass)."""
import numpy as np
import pytest

import tensum_amd as T
from tensum_amd import uq

pytestmark = pytest.mark.gpu


def _levels(problem, meshes, counts, seed):
    out = []
    for n, k in zip(meshes, counts):
        inp = uq.sample_inputs(k, problem.bounds, seed=seed + n)
        out.append(uq.make_batch(problem, problem.mesh(n), inp))
    return out


@pytest.mark.parametrize("name", ["vortex", "riemann"])
def test_packed_steps_equal_separate_runs(name):
    print(f"packed {name}", flush=True)
    p = uq.vortex_problem(final_time=0.4) if name == "vortex" else uq.riemann_problem(final_time=0.01)
    meshes, counts = (17, 33, 65), (16, 4, 2)
    packed = _levels(p, meshes, counts, seed=5)
    alone = _levels(p, meshes, counts, seed=5)
    for chunk in (7, 50, 10 ** 4):
        done = T.run_packed(packed, chunk)
        for a, d in zip(alone, done):
            assert int(np.sum(a.run(chunk))) == int(d)
        for a, b in zip(packed, alone):
            for k in range(a.n_samples):
                assert np.array_equal(a.get_state(k), b.get_state(k))
                assert a.time(k) == b.time(k)
    for a in packed:  # every sample reached its final time
        for k in range(a.n_samples):
            assert a.time(k)[0] == p.params.final_time


def test_packed_refuses_unsupported_contexts():
    p = uq.vortex_problem(final_time=0.1)
    a = _levels(p, (17,), (2,), seed=1)[0]
    m = p.mesh(17)
    ns = T.Solver(m, T.Params(model="ns", flux="kepec", mu_ref=1e-3, final_time=0.1))
    with pytest.raises(T.TsmError, match="packed"):
        T.run_packed([a, ns], 5)
    other = T.Solver(m, T.Params(flux="kepes_tecno_rusanov", reconstruct="minmod", final_time=0.1))
    with pytest.raises(T.TsmError, match="share"):
        T.run_packed([a, other], 5)
