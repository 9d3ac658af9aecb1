"""Writes tests/golden/ref_riemann1025_stop.npz: where the reference's own run
of the 1025^2 four-quadrant Riemann problem (KEPES-TeCNO Rusanov, minmod,
SSP-RK3, CFL 0.5, seed-1 jittered mesh) ends when asked for t_final 0.8.

The scheme is not positivity preserving: on this mesh the residual norm grows
without bound after some hundred steps and turns NaN, and the reference ends
its time loop there (`NaN > min_residue` is false, src/fv.cc:2273).  The
fixture keeps the number of steps it took and the iteration, dt and t of its
last 64 steps (residual.dat).

The reference (oracle/_ref/tensum and its partitioner oracle/_ref/grid_part,
built from its sources by oracle/ref.mk) runs on RANKS MPI ranks (default 8)
on the mesh split into x-strips (cases.write_msh_strips), as bench.py's CPU
baseline runs it; grid_part re-prints the coordinates at 15-16 significant
digits.  It needs the reference's binaries and MPICH, so it runs only in the
container that holds the reference.

Usage: python tests/golden/make_ref_riemann1025_stop.py [RANKS]
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests"), HERE]

import bench  # noqa: E402
from cases import riemann_mesh  # noqa: E402

OUT = os.path.join(HERE, "ref_riemann1025_stop.npz")


def main():
    ranks = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    mesh = riemann_mesh(1025)
    with tempfile.TemporaryDirectory() as d:
        r = bench._reference_run(mesh, 0.8, ranks, d, lambda m: print(m, flush=True))
        if r is None:
            raise SystemExit("the reference run failed")
        rows = [ln.split() for ln in open(os.path.join(d, "SAMPLE_0", "residual.dat"))
                if ln.strip()]
    steps = len(rows)
    last = np.array([[float(x) for x in row[:3]] for row in rows[-64:]])
    final_res = rows[-1][3]
    np.savez(OUT, steps=steps, iter_dt_t=last, ranks=ranks, final_res=final_res)
    print(f"{OUT}: {steps} steps, last t {last[-1, 2]!r}, last residual {final_res}")


if __name__ == "__main__":
    main()
