"""The C-ABI library loads and exports every entry point include/boltx.h
declares; argument checks reject bad input before any device work."""
import ctypes
import os
import re

from bolt_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "boltx.h")).read()
    return sorted(set(re.findall(r"\b(boltx_[a-z_]+)\s*\(", text)))


def test_exports_match_header():
    lib = _lib.load()
    names = declared()
    assert set(names) == set(_lib.EXPORTS), names
    for n in names:
        assert hasattr(lib, n)
    assert lib.boltx_version() == 1


def test_invalid_arguments_are_rejected():
    lib = _lib.load()
    rc = lib.boltx_fvm_stage(None, None, None, None)
    assert rc == -1
    assert b"grid" in lib.boltx_last_error()
    g = _lib.Grid(4, 4, 8, 8, 1, 2, 0.25, 0.25, 1.0, 1.0, 1.0, -4.0, -4.0, -0.5)
    rc = lib.boltx_moments(ctypes.byref(g), 0, ctypes.c_void_p(1), ctypes.c_void_p(1), None)
    assert rc == -3 and b"species" in lib.boltx_last_error()
    g.n_species = 1
    ph = _lib.Physics(3, 3, 0, 0, 1, 0, 2, 7, 1.0, 1.0, 1.0, 0.01)
    st = _lib.Stage()
    st.u = st.w = st.out = ctypes.c_void_p(8)
    rc = lib.boltx_fvm_stage(ctypes.byref(g), ctypes.byref(ph), ctypes.byref(st), None)
    assert rc == -1 and b"precision" in lib.boltx_last_error()
    ph.precision = 0
    rc = lib.boltx_fvm_stage(ctypes.byref(g), ctypes.byref(ph), ctypes.byref(st), None)
    assert rc == -1 and b"alias" in lib.boltx_last_error()
