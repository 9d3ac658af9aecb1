"""The C-ABI library loads without a GPU and exports every function that
include/tsm.h declares."""
import ctypes as C
import os
import re

import tensum_amd as T

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "tsm.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tsm_\w+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared()
    for n in ("tsm_create", "tsm_run", "tsm_residual", "tsm_moments", "tsm_sobol",
              "tsm_mlp_train", "tsm_run_group"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(T.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
