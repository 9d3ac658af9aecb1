"""The oracle against the reference solver's own outputs (CPU).

The fixtures tests/golden/ref_<case>.npz were written by the reference
binary (built from its own sources by oracle/ref.mk, run by
tests/golden/make_ref_fixtures.py) on this build's seeded meshes.  They pin
the oracle where no closed form does: the residual-norm convention of
residual.dat (src/fv.cc:1801-1879), the boundary states and boundary fluxes
of slip, outlet, inlet, farfield and pressure sides (src/bc.h:569-654,
src/material.cc:110-127), the Liou fix with eigenvalue modification
(src/compute_ssw.cc:7-39), all six two-point fluxes (KEP's and simple_avg's
included), the first, second, minmod, TVB, ENO2 and van Albada
reconstructions, Voronoi duals with RK4, the viscous residual with adiabatic
and isothermal noslip wall cells, the error norms of error_norms.dat and the
integrals of global.dat.
"""
import pytest

import oracle_binding as O
from ref_fixture_check import check_case

FAST = ["riemann33", "vortex33", "mixed23", "pressure23", "liou25", "vortex17_rk4_voronoi",
        "tgv33_ns", "kep_vortex17", "roe_ec_vortex17", "avg_vortex17", "albada_vortex25",
        "second_vortex25", "eno2_riemann25", "first_riemann25", "walls_ns23"]


@pytest.mark.parametrize("case", FAST)
def test_oracle_matches_reference_run(case):
    check_case(O.Oracle, case)


@pytest.mark.slow_cpu
def test_oracle_matches_reference_vortex_129_to_t10():
    """The 129^2 isentropic vortex over one period (2000+ steps): the whole
    residual log, the final state and the error norms."""
    check_case(O.Oracle, "vortex129_t10")
