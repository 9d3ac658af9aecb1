"""Comparison of a solver run with a fixture the reference solver wrote.

Shared by the oracle's CPU tests and libtsm's GPU tests: `solver` is either
oracle_binding.Oracle or tensum_amd.Solver (the same method names).  The
fixtures (tests/golden/ref_<case>.npz, written by
tests/golden/make_ref_fixtures.py) hold the reference's residual.dat,
the restart file after a kept step, and, per case, error_norms.dat and
global.dat.

Tolerances: the restart state is printed at 16 significant digits, so
conserved variables agree to rel-L2 1e-10 (the parity bar) with room for the
ulp-level differences of the fused multiply-adds and GPU logarithms; dt, t and
the residual norm are printed at 12 digits (src/fv.cc:1959-1968); error norms
at 9 (src/fv.cc:1758-1790); global.dat at 6 (src/fv.cc:2218-2220).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from cases import conserved, rel_l2  # noqa: E402
from ref_cases import CASES, params  # noqa: E402


def fixture(case):
    return np.load(os.path.join(HERE, "golden", f"ref_{case}.npz"))


def check_case(make_solver, case, **extra):
    """Runs `case` to its kept step (and on to its final time when the
    fixture holds error norms) and compares with the fixture."""
    c = CASES[case]
    f = fixture(case)
    mesh = c["mesh"]()
    prm = params(case, float(f["final_time"]))
    for k, v in extra.items():
        setattr(prm, k, v)
    s = make_solver(mesh, prm)
    s.set_state(c["state"](mesh))
    ref = f["log"]
    if c["steps"]:
        n, log = s.run(c["steps"], want_log=True)
        n = int(np.asarray(n).ravel()[0])
        assert n == c["steps"]
        assert rel_l2(conserved(s.get_state()), conserved(f["state"])) <= 1e-10
        it, counter, t, res0 = s.clock()
        assert it == int(f["clock"][0]) and counter == int(f["clock"][1])
        assert abs(t - f["clock"][2]) <= 1e-12 * abs(f["clock"][2])
        assert abs(res0 - f["clock"][3]) <= 1e-9 * abs(f["clock"][3])
        m = c["steps"]
    else:
        n, log = s.run(ref.shape[0] + 5, want_log=True)
        n = int(np.asarray(n).ravel()[0])
        assert n == ref.shape[0]
        assert rel_l2(conserved(s.get_state()), conserved(f["state"])) <= 1e-10
        m = n
    log = np.asarray(log).reshape(-1, 4)[:m]
    np.testing.assert_array_equal(log[:, 0], ref[:m, 0])
    np.testing.assert_allclose(log[:, 1], ref[:m, 1], rtol=1e-11, atol=0)   # dt
    np.testing.assert_allclose(log[:, 2], ref[:m, 2], rtol=1e-11, atol=0)   # t
    np.testing.assert_allclose(log[:, 3], ref[:m, 3], rtol=1e-9, atol=0)    # res_total
    if "glob" in f.files:  # kinetic energy and entropy after the last step compared
        ke, ent = s.global_integrals()
        g = f["glob"][m - 1]
        assert int(g[0]) == m
        np.testing.assert_allclose([ke, ent], g[2:4], rtol=6e-6)
    if "error" in f.files:
        if c["steps"]:  # on to the final time of the reference's run
            n2, _ = s.run(ref.shape[0] - m + 5, want_log=True)
            assert m + int(np.asarray(n2).ravel()[0]) == ref.shape[0]
        t = s.time()
        t = t[0] if isinstance(t, tuple) else t
        assert abs(t - ref[-1, 2]) <= 1e-11 * ref[-1, 2]
        exact = c["exact_state"](mesh, t) if "exact_state" in c else c["state"](mesh)
        err = np.asarray(s.error_norms(exact)).reshape(8, 3)
        np.testing.assert_allclose(err, f["error"], rtol=2e-8, atol=1e-300)
    return s
