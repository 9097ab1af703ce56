"""CPU-side checks of the drop-in boundary: the C-ABI library loads, exports
every entry point include/vespahip.h declares, and reports errors (not crashes)
when no device is present. No compute calls are made here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def header_functions():
    txt = open(os.path.join(ROOT, "include", "vespahip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(vh_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_what_the_binding_lists(V):
    assert header_functions() == sorted(V.EXPORTED)


def test_library_exports_every_declared_symbol(V):
    L = V.load_library()
    missing = [s for s in header_functions() if not hasattr(L, s)]
    assert not missing


def test_abi_version_and_no_error_initially(V):
    L = V.load_library()
    assert L.vh_abi_version() == 1
    assert V.device_count() >= 0


def test_missing_library_fails_loudly(V, tmp_path):
    with pytest.raises(ImportError, match="no CPU fallback"):
        V.load_library(str(tmp_path / "libvespahip.so"))


def test_argument_validation_without_device(V):
    """Host-side validation runs before any device call."""
    L = V.load_library()
    h = C.c_void_p()
    dims = (C.c_int64 * 3)(8, 8, 8)
    lo = (C.c_double * 3)(0, 0, 0)
    hi = (C.c_double * 3)(1, 1, 1)
    bc = (C.c_int32 * 6)(0, 1, 0, 0, 0, 0)   # periodic on one side only
    assert L.vh_grid_create(C.byref(h), 3, dims, lo, hi, bc, 1.4, 0) != 0
    assert b"periodic" in L.vh_last_error()
    bc = (C.c_int32 * 6)(0, 0, 0, 0, 0, 0)
    assert L.vh_grid_create(C.byref(h), 4, dims, lo, hi, bc, 1.4, 0) != 0
    assert b"rank" in L.vh_last_error()
    assert L.vh_grid_create(C.byref(h), 3, dims, lo, hi, bc, 1.0, 0) != 0
    assert b"gamma" in L.vh_last_error()
    assert L.vh_face_flux_f64(None, None, None, -1, 1.4, None) != 0
    assert L.vh_face_flux_f64(None, None, None, 0, 1.4, None) == 0      # nf == 0 no-op
    # axis outside 1..3 is rejected before any copy (the reference's tuple index raises)
    W = np.ones((3, 5))
    F = np.zeros((3, 5))
    dp = C.POINTER(C.c_double)
    for bad in (0, 4, -1):
        ax = np.array([1, bad, 2], dtype=np.int32)
        rc = L.vh_face_flux_f64(W.ctypes.data_as(dp), W.ctypes.data_as(dp),
                                ax.ctypes.data_as(C.POINTER(C.c_int32)), 3, 1.4, F.ctypes.data_as(dp))
        assert rc == -1 and b"axis" in L.vh_last_error()


def test_host_mirror_validation(V):
    with pytest.raises(ValueError, match="γ must exceed 1"):
        V.Problem("x", (4,), ((0, 1),), 1.0, "outflow", None, 1.0)
    with pytest.raises(ValueError, match="cfl"):
        V.Problem("x", (4,), ((0, 1),), 1.4, "outflow", None, 1.0, cfl=1.5)
    with pytest.raises(ValueError, match="hi > lo"):
        V.Problem("x", (4,), ((1, 0),), 1.4, "outflow", None, 1.0)
    with pytest.raises(TypeError):
        V.compute_face_fluxes(object(), V.IdealHydro(1.4), np.zeros((1, 5)), np.zeros((1, 5)),
                              np.zeros((1, 5)), np.ones(1, dtype=np.int32))


def test_product_does_not_import_the_oracle():
    """The oracle is the checker only: nothing under vespa.jl_amd/ may reference it."""
    pkg = os.path.join(ROOT, "vespa.jl_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp", ".jl", "Makefile")):
                txt = open(os.path.join(dirpath, f), errors="replace").read()
                assert "oracle" not in txt.lower(), os.path.join(dirpath, f)
