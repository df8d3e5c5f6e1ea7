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


def test_initialiser_and_user_term_argument_checks():
    """The device-initialiser and user-term entry points reject bad
    arguments before any device work."""
    from fgpe_amd import _lib
    L = _lib.load()
    g = _lib.Geom.of((1.0, 1.0, 1.0), (0.1, 0.1, 0.1))
    ax = (ctypes.c_double * 3)(1.0, 1.0, 1.0)
    assert L.fgpe_psi_init(None, 0, 0, ctypes.byref(g), ax, 0.0, 0.0, 1) == -1
    assert L.fgpe_psi_normalize(None, 0, 1.0, ctypes.byref(g), None) == -1
    assert L.fgpe_vext_init(None, 0, 0, ctypes.byref(g), ax) == -1
    assert L.fgpe_vext_commit(None, 0, None) == -1
    assert L.fgpe_user_term(None) == -1
    v = ctypes.c_int(7)
    assert L.fgpe_user_term(ctypes.byref(v)) == 0
    assert v.value == 0   # the shipped library compiles no user term
