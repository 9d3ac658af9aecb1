"""Writes tests/golden/ref_<case>.npz by running the reference solver.

The reference binary is oracle/_ref/tensum, built from the reference's own
sources by oracle/ref.mk (make -C oracle -f ref.mk); it runs only in the
container that holds the reference.  For every case of ref_cases.py this
script writes the part file of the build's seeded mesh and a param.in, runs
the reference (NPART = 1), and keeps the outputs listed in ref_cases.py.
The oracle is used only to choose a final time a little past the step whose
restart file is kept (unsteady runs stop only at final_time).

Usage: python tests/golden/make_ref_fixtures.py [case ...]
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests"), HERE]

import oracle_binding as O  # noqa: E402
from cases import write_part_file  # noqa: E402
from ref_cases import CASES, param_text, params, periodic_tags  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "tensum")


def _stop_time(case, mesh):
    """A final time two and a half steps past the time of step `steps`, so the
    run writes no second restart file into head 0 after that step."""
    c = CASES[case]
    o = O.Oracle(mesh, params(case, 1e6))
    o.set_state(c["state"](mesh))
    n, log = o.run(c["steps"], want_log=True)
    assert n == c["steps"]
    return float(f"{log[-1, 2] + 2.5 * log[-1, 1]:.6g}")


def _error_norms(path):
    rows = []
    for line in open(path):
        if "(L1, L2, Linf):" in line:
            rows.append([float(v) for v in line.split(":")[1].split()])
    return np.array(rows)


def run_case(case):
    c = CASES[case]
    mesh = c["mesh"]()
    final_time = c["final_time"] if c["final_time"] else _stop_time(case, mesh)
    freq = c["steps"] if c["steps"] else 10 ** 8
    with tempfile.TemporaryDirectory() as d:
        write_part_file(mesh, d, periodic_tags=periodic_tags(case))
        with open(os.path.join(d, "param.in"), "w") as f:
            f.write(param_text(case, final_time, freq))
        r = subprocess.run([EXE, "-i", "param.in", "-npart", "1", "-L", "."], cwd=d,
                           capture_output=True, text=True, timeout=3600)
        sd = os.path.join(d, "SAMPLE_0")
        if r.returncode != 0 or not os.path.exists(os.path.join(sd, "residual.dat")):
            raise RuntimeError(f"{case}: reference run failed\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        log = np.loadtxt(os.path.join(sd, "residual.dat"), ndmin=2)
        out = {"log": log, "final_time": np.array(final_time)}
        # restart head 0 holds step `steps` (the first write), or the final
        # state when the case keeps no intermediate step
        head = 0 if c["steps"] else int(open(os.path.join(sd, "LAST_RESTART_HEAD.txt")).read())
        rst = open(os.path.join(sd, f"restart_{head}_0.dat")).read().split()
        nv = mesh.n_vertex
        vals = np.array([float(v) for v in rst[:5 * nv]]).reshape(nv, 5)
        out["state"] = vals[:, [0, 1, 2, 4]]
        out["clock"] = np.array([float(v) for v in rst[5 * nv:5 * nv + 4]])
        if c["steps"]:
            assert int(out["clock"][0]) == c["steps"], out["clock"]
            assert log.shape[0] > c["steps"], "the run stopped before the kept step"
        if c.get("exact"):
            out["error"] = _error_norms(os.path.join(sd, "error_norms.dat"))
        if c.get("save_global"):
            out["glob"] = np.loadtxt(os.path.join(sd, "global.dat"), ndmin=2)
    np.savez_compressed(os.path.join(HERE, f"ref_{case}.npz"), **out)
    print(f"{case}: {log.shape[0]} steps to t = {log[-1, 2]:.6g}, kept step "
          f"{int(out['clock'][0])}", flush=True)


# The output files themselves, for the writer's format: the 33^2 vortex to
# t = 0.2 with two solution stamps, error norms, global integrals and a restart.
OUT_CASE, OUT_TIME = "vortex33", 0.2
OUT_FILES = ("residual.dat", "error_norms.dat", "global.dat", "master_file.visit",
             "LAST_RESTART_HEAD.txt", "restart_0_0.dat", "sol0000_0.vtk", "sol0001_0.vtk",
             "sol0002_0.vtk")


def output_param_text():
    """param.in of the output fixture (also used by the test that runs the
    driver on the same case)."""
    t = param_text(OUT_CASE, OUT_TIME, 10 ** 8)
    for a, b in (("   soln_output           no", "   soln_output           yes"),
                 ("   time_stamps           1", "   time_stamps           2"),
                 ("   save_global           no", "   save_global           yes")):
        assert a in t
        t = t.replace(a, b)
    return t


def run_output_case():
    mesh = CASES[OUT_CASE]["mesh"]()
    with tempfile.TemporaryDirectory() as d:
        write_part_file(mesh, d, periodic_tags=periodic_tags(OUT_CASE))
        with open(os.path.join(d, "param.in"), "w") as f:
            f.write(output_param_text())
        r = subprocess.run([EXE, "-i", "param.in", "-npart", "1", "-L", "."], cwd=d,
                           capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"output case: reference run failed\n{r.stdout[-2000:]}")
        files = {}
        for name in OUT_FILES:
            with open(os.path.join(d, "SAMPLE_0", name), "rb") as f:
                files[name.replace(".", "_")] = np.frombuffer(f.read(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, f"ref_out_{OUT_CASE}.npz"), **files)
    print(f"output files of {OUT_CASE} to t = {OUT_TIME}: {', '.join(OUT_FILES)}", flush=True)


if __name__ == "__main__":
    if not os.path.exists(EXE):
        sys.exit("build the reference first: make -C oracle -f ref.mk")
    for case in (sys.argv[1:] or list(CASES) + ["outputs"]):
        if case == "outputs":
            run_output_case()
        else:
            run_case(case)
