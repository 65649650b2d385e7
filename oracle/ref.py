"""ctypes front end of oracle/_ref/libefdc_ref.so (the reference's own
CALTRAN/CALFQC and SETOPENBC/SETOPENBC2 compiled from its sources).

TEST INFRASTRUCTURE: used only by tests to pin the C restatement.  The
library exists only where the reference sources were present at build
time (oracle/build_ref.py).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libefdc_ref.so")

_lib = None


def available() -> bool:
    return os.path.exists(LIB)


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(LIB)
        i, d, p, i64 = C.c_int, C.c_double, C.c_void_p, C.c_int64
        _lib.ref_init.argtypes = [i] * 15
        _lib.ref_scalars.argtypes = [d, d, d, d, i, i]
        _lib.ref_flags.argtypes = [i, i, i]
        for f in ("ref_put_r", "ref_put_i", "ref_put_l8", "ref_get_r",
                  "ref_get_i", "ref_get_l8"):
            getattr(_lib, f).argtypes = [C.c_char_p, p, i64]
        _lib.ref_caltran.argtypes = [i, i, p, p, d, i64]
        _lib.ref_setopenbc.argtypes = [d, i64]
        _lib.ref_wetlist.argtypes = [p]
    return _lib


def put(name: str, arr: np.ndarray):
    """Copy a flat host array, already in Fortran element order, into the
    GLOBAL array of that name (var_global layout arrays of a Case are in
    that order when raveled; use fortran() for arrays built with their
    Fortran shape)."""
    a = np.ascontiguousarray(np.asarray(arr).ravel())
    if a.dtype == np.float64:
        f = lib().ref_put_r
    elif a.dtype == np.int64:
        f = lib().ref_put_l8
    else:
        a = np.ascontiguousarray(a.astype(np.int32))
        f = lib().ref_put_i
    f(name.encode(), a.ctypes.data, a.size)


def fortran(a: np.ndarray) -> np.ndarray:
    """Flatten an array given in its Fortran shape into Fortran order."""
    return np.asarray(a).reshape(-1, order="F")


def get(name: str, n: int, dtype=np.float64) -> np.ndarray:
    a = np.zeros(n, dtype=dtype)
    f = {np.float64: lib().ref_get_r, np.int32: lib().ref_get_i,
         np.int64: lib().ref_get_l8}[dtype]
    f(name.encode(), a.ctypes.data, n)
    return a


MVAR_OF_MO = {1: 1, 2: 2, 3: 3}


def mvar_of(mo: int) -> int:
    return MVAR_OF_MO.get(mo, 6)
