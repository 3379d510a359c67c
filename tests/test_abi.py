"""The C-ABI library loads and exports every entry point include/ccd_amd.h declares (no GPU needed)."""
import ctypes
import os
import re

from collisiondetection_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ccd_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*\*?\s*(ccd_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_native.LIB_PATH), "build libccd_amd.so first (__graft_entry__.build())"
    L = ctypes.CDLL(_native.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 12
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_native.SIGNATURES), set(syms) ^ set(_native.SIGNATURES)


def test_binding_signatures_load():
    L = _native.lib()
    assert L.ccd_last_error(None) == b"null context"
    assert L.ccd_destroy(None) == _native.CCD_E_INVALID
