"""The C-ABI library loads and exports every entry point include/ctwin.h
declares (CPU only: no compute calls)."""
import ctypes
import os
import re

import pytest

import ctwin_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "ctwin.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ct_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "ct_engine_create" in names and "ct_sieve_window" in names and "ct_prp_check" in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(ctwin_amd.LIB_PATH):
        pytest.skip("libctwin_amd.so not built (run python -m ctwin_amd.build)")
    L = ctypes.CDLL(ctwin_amd.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    assert set(ctwin_amd.EXPORTS) <= set(declared())


def test_errors_without_device_are_status_codes():
    if not os.path.exists(ctwin_amd.LIB_PATH):
        pytest.skip("libctwin_amd.so not built")
    L = ctwin_amd.lib()
    h = ctypes.c_void_p()
    # invalid sizes are rejected before any device call
    assert L.ct_engine_create(0, 40, 1, 16, 0, ctypes.byref(h)) == -4
    assert L.ct_engine_create(0, 10, 0, 16, 0, ctypes.byref(h)) == -1
    assert L.ct_engine_create(0, 10, 1, 8, 0, ctypes.byref(h)) == -4
    assert b"log2_n" in L.ct_last_error() or len(L.ct_last_error()) > 0
