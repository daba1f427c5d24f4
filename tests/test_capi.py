"""The C-ABI library loads and exports every symbol include/hm_stitch.h
declares (no compute call, so this runs without a GPU)."""
import ctypes
import os

import pytest


def test_header_symbols_exported():
    import hm_stitch

    if not os.path.exists(hm_stitch.LIB_PATH):
        pytest.fail("libhm_stitch.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(hm_stitch.LIB_PATH)
    names = hm_stitch.exported_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_signatures_cover_header():
    import hm_stitch

    lib = hm_stitch.lib()
    for n in hm_stitch.exported_symbols():
        assert getattr(lib, n).argtypes is not None, f"{n} has no ctypes signature"


def test_oracle_is_not_linked_by_the_product():
    import hm_stitch

    with open(hm_stitch.LIB_PATH, "rb") as f:
        blob = f.read()
    assert b"hmo_" not in blob and b"hm_oracle" not in blob
