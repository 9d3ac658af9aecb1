"""Autotuner host logic (no GPU): candidate launch shapes and the cache."""
import json

import tensum_amd as T
from tensum_amd import autotune as AT


def test_candidates_respect_tile_and_block():
    c = AT.candidates(tiles=(64, 256), fu_blocks=(0, 128, 512), grad_blocks=(0,))
    assert {"tile_vertices": 256, "block_fu": 128, "block_grad": 0, "fused_gradient": False} not in c
    assert {"tile_vertices": 64, "block_fu": 128, "block_grad": 0, "fused_gradient": False} in c
    sep = AT.candidates(tiles=(128,), fu_blocks=(0,), grad_blocks=(0, 256), fused=(False, True))
    assert len(sep) == 3  # the gradient block is varied only for the separate pass


def test_cache_stores_and_reuses(tmp_path):
    path = str(tmp_path / "tune.json")
    calls = []

    def tune():
        calls.append(1)
        return {"tile_vertices": 96, "block_fu": 0, "block_grad": 0, "fused_gradient": False}

    a = AT.cached(path, "k", tune)
    b = AT.cached(path, "k", tune)
    assert a == b and len(calls) == 1
    assert json.load(open(path))["k"]["tile_vertices"] == 96
    p = AT.apply(T.Params(), a)
    assert AT.shape_of(p) == a
