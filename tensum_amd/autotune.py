"""Launch-shape autotuner for the tiled stage kernels.

Chooses the vertex tile size (owned vertices per workgroup, which sets the
halo fraction and the faces each tile recomputes) and the workgroup sizes of
the stage and gradient kernels, and whether the gradients are formed inside
the stage kernel (two halo rings) or in a pass of their own, by timing a few steps of the actual workload
with the solver's own event timer (tsm_bench).  Every candidate computes the
same sums in the same order, so tuning changes only the time: the tests check
that a tuned run is bit-identical to an untuned one.  The reference has no
counterpart (its loop is one CPU thread per rank, src/fv.cc:2273-2368).
"""
from __future__ import annotations

import dataclasses
import json
import os

from . import Params, Solver, TsmError

TILE_SIZES = (64, 96, 128, 160, 192, 256)
FU_BLOCKS = (0, 256, 512)      # 0: the smallest block that covers the tile
GRAD_BLOCKS = (0, 256)


def candidates(tiles=TILE_SIZES, fu_blocks=FU_BLOCKS, grad_blocks=GRAD_BLOCKS, separate=True,
               fused=(False,)):
    """Launch shapes to try; the gradient block only matters for the separate
    gradient pass, so it is varied only there."""
    out = []
    for fu in fused:
        for tb in tiles:
            for bf in fu_blocks:
                if bf and bf < tb:
                    continue
                for bg in (grad_blocks if (separate and not fu) else (0,)):
                    if bg and bg < tb:
                        continue
                    out.append({"tile_vertices": tb, "block_fu": bf, "block_grad": bg,
                                "fused_gradient": fu})
    return out


def autotune(mesh, params: Params, state, n_samples=1, steps=10, shapes=None, log=None):
    """Times every launch shape on `mesh` with `state` in every sample and
    returns (best Params, table of (shape, step_us)).  Shapes the kernels
    cannot take (LDS or block limits) are skipped."""
    shapes = shapes if shapes is not None else candidates(separate=not params.fused_gradient)
    table = []
    for sh in shapes:
        prm = dataclasses.replace(params, **sh)
        try:
            s = Solver(mesh, prm, n_samples=n_samples)
        except TsmError:
            continue
        for k in range(n_samples):
            s.set_state(state, k)
        s.run(1)  # first step (and the lazy set-up) outside the timing
        us = s.bench(steps)["step_us"]
        s.close()
        table.append((sh, us))
        if log:
            log(f"autotune {sh}: {us:.1f} us/step")
    if not table:
        raise TsmError("no launch shape fits this mesh")
    best = min(table, key=lambda r: r[1])[0]
    return dataclasses.replace(params, **best), table


def cached(path, key, tune):
    """Returns the shape stored under `key` in the JSON file `path`, or runs
    tune() -> shape dict and stores it."""
    db = {}
    if os.path.exists(path):
        with open(path) as f:
            db = json.load(f)
    if key not in db:
        db[key] = tune()
        with open(path, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
    return db[key]


def workload_key(mesh, params: Params, n_samples):
    return (f"nv{mesh.n_vertex}-nb{mesh.bface_tag.size}-np{max(mesh.periodic_ptr.size - 1, 0)}-"
            f"{params.flux}-{params.reconstruct}-{params.model}-S{n_samples}")


def shape_of(params: Params):
    return {"tile_vertices": params.tile_vertices, "block_fu": params.block_fu,
            "block_grad": params.block_grad, "fused_gradient": params.fused_gradient}


def apply(params: Params, shape):
    return dataclasses.replace(params, **shape)
