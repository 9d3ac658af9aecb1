"""libtsm (HIP, gfx950) against the reference solver's own outputs.

The same fixtures as tests/test_ref_fixtures.py (the reference binary's
residual.dat, restart file, error_norms.dat and global.dat on this build's
meshes), compared with the GPU runs at the same tolerances: conserved
variables to rel-L2 1e-10 after the kept step, dt and t to 1e-11, the
residual norm to 1e-9, error norms to 2e-8 (printed at 9 digits), integrals
to the 6 digits of global.dat.  Each case runs on the kernels its options
select by default (the LDS-tiled stage kernels, or the CSR-gather kernels
for the Liou fix) and again on the CSR-gather kernels.
"""
import pytest

import tensum_amd as T
from ref_fixture_check import check_case

pytestmark = pytest.mark.gpu

FAST = ["riemann33", "vortex33", "mixed23", "pressure23", "liou25", "vortex17_rk4_voronoi",
        "tgv33_ns", "kep_vortex17", "roe_ec_vortex17", "avg_vortex17", "albada_vortex25",
        "second_vortex25", "eno2_riemann25", "first_riemann25", "walls_ns23"]


@pytest.mark.parametrize("path", ["auto", "csr"])
@pytest.mark.parametrize("case", FAST)
def test_gpu_matches_reference_run(case, path):
    print(f"reference fixture {case} on the {path} kernels", flush=True)
    check_case(T.Solver, case, kernel_path=path)


@pytest.mark.parametrize("graph", [0, 8])
def test_gpu_matches_reference_vortex_129_to_t10(graph):
    """The plain-method workload: 129^2 isentropic vortex, KEPES-TeCNO Roe,
    minmod, SSP-RK3, CFL 0.8 over one period (t = 10, 2000+ steps): every
    step's dt, t and residual norm, the final state and the error norms of
    the reference's run (far inside the 3 significant figures asked for)."""
    print(f"reference fixture vortex129_t10, graph_steps {graph}", flush=True)
    check_case(T.Solver, "vortex129_t10", graph_steps=graph)
