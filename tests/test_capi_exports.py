"""The C-ABI library loads and exports every symbol include/fgpe_c.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdrs = [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))
            if f.endswith(".h")]
    names = set()
    for h in hdrs:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(fgpe_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "fgpe_rk4_steps" in names and "fgpe_ctx_create" in names
    from fgpe_amd import _lib
    assert set(names) == set(_lib.EXPORTS)


def test_library_exports_all_declared_symbols():
    from fgpe_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libfgpe.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_error_path_without_device_work():
    from fgpe_amd import _lib
    L = _lib.load()
    h = ctypes.c_void_p()
    # invalid precision is rejected before any device call
    assert L.fgpe_ctx_create(64, 64, 64, 7, ctypes.byref(h)) == -1
    assert b"precision" in L.fgpe_last_error()
    assert L.fgpe_ctx_create(63, 64, 64, 0, ctypes.byref(h)) == -2
    assert L.fgpe_ctx_destroy(None) == 0
