"""The allocation-free tier-1 entry point (vh_face_flux_noalloc) against the
allocating host-pointer form (vh_face_flux_f64/_f32) and the oracle, through the
C ABI: two calls in a row on one device (the second smaller, so stale staging
from the first would show), for both element types, plus the size limit."""
import ctypes as C
import sys

import numpy as np
import pytest

from helpers import random_faces

LIMIT = 262144


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _noalloc(L, WL, WR, ax, gamma, dtype_code):
    F = np.full_like(WL, np.nan)
    rc = L.vh_face_flux_noalloc(WL.ctypes.data, WR.ctypes.data, ax.ctypes.data_as(C.POINTER(C.c_int32)),
                                len(ax), gamma, dtype_code, F.ctypes.data, None)
    return rc, F


def _alloc(L, WL, WR, ax, gamma, dtype):
    F = np.full_like(WL, np.nan)
    tp = C.POINTER(C.c_double if dtype == np.float64 else C.c_float)
    f = L.vh_face_flux_f64 if dtype == np.float64 else L.vh_face_flux_f32
    rc = f(WL.ctypes.data_as(tp), WR.ctypes.data_as(tp), ax.ctypes.data_as(C.POINTER(C.c_int32)), len(ax),
           gamma, F.ctypes.data_as(tp))
    return rc, F


def test_noalloc_rejects_oversized_and_bad_input(V):
    """Host-side checks run before any device work (no GPU needed)."""
    L = V.load_library()
    WL = np.ones((LIMIT + 1, 5))
    ax = np.ones(LIMIT + 1, dtype=np.int32)
    rc, _ = _noalloc(L, WL, WL, ax, 1.4, 0)
    assert rc == -1 and b"VH_NOALLOC_MAX_FACES" in L.vh_last_error()
    rc, _ = _noalloc(L, WL[:4], WL[:4], np.array([1, 2, 0, 3], dtype=np.int32), 1.4, 0)
    assert rc == -1 and b"axis" in L.vh_last_error()
    rc, _ = _noalloc(L, WL[:4], WL[:4], ax[:4], 1.4, 7)
    assert rc == -1 and b"dtype" in L.vh_last_error()
    assert L.vh_face_flux_noalloc(None, None, None, 0, 1.4, 0, None, None) == 0     # nf == 0 no-op


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,code", [(np.float64, 0), (np.float32, 1)])
def test_noalloc_matches_allocating_form_twice(V, O, dtype, code):
    L = V.load_library()
    for call, nf in enumerate((LIMIT, 1000)):
        _log(f"noalloc dtype={np.dtype(dtype).name} call={call} nf={nf}")
        WL, WR, ax = random_faces(nf, seed=100 + call, dtype=dtype)
        rc, F = _noalloc(L, WL, WR, ax, 1.4, code)
        assert rc == 0, L.vh_last_error()
        rc, Fa = _alloc(L, WL, WR, ax, 1.4, dtype)
        assert rc == 0, L.vh_last_error()
        assert np.array_equal(F, Fa)
        assert np.array_equal(F, O.face_flux(WL, WR, ax, 1.4))
