"""The C-ABI library loads and exports every entry point include/pisces_hip.h
declares (no compute calls: this runs without a GPU)."""
import ctypes

import pytest


def test_library_exports_header_symbols(pisces):
    names = pisces.exported_symbols()
    assert "pisces_solve" in names and "pisces_apply_h" in names
    lib = pisces.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_errors(pisces):
    lib = pisces.lib()
    assert b"gfx950" in lib.pisces_version()
    assert lib.pisces_strerror(-1) == b"invalid argument"
    assert lib.pisces_strerror(3) == b"not converged"


def test_argument_checks_need_no_gpu(pisces):
    lib = pisces.lib()
    h = ctypes.c_void_p()
    import numpy as np
    n = np.array([1, 8, 8], dtype=np.int32)
    L = np.array([1.0, 1.0, 1.0])
    rc = lib.pisces_create(ctypes.byref(h), n.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                           L.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 0)
    assert rc == -1
