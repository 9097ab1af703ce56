"""ctypes binding of libvespahip.so (the C ABI declared in include/vespahip.h).

The library is built in-tree by ``vespa.jl_amd/build.py`` (hipcc, gfx950). If it
is missing this module raises: there is no CPU fallback for the product path.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# VESPAHIP_LIB overrides the library location (e.g. an alternative build).
LIB_PATH = os.environ.get("VESPAHIP_LIB") or os.path.join(PKG_DIR, "libvespahip.so")

VH_F64, VH_F32 = 0, 1
VH_EXACT, VH_FAST = 0, 1
BC_PERIODIC, BC_OUTFLOW, BC_REFLECTING, BC_PARENT = 0, 1, 2, 3

# Every symbol include/vespahip.h declares (checked by the CPU test suite).
EXPORTED = (
    "vh_last_error", "vh_abi_version", "vh_device_count", "vh_device_cu_count",
    "vh_face_flux_f64", "vh_face_flux_f32", "vh_face_flux_dev_f64", "vh_face_flux_dev_f32",
    "vh_face_flux_noalloc",
    "vh_buffer_alloc", "vh_buffer_free", "vh_buffer_upload", "vh_buffer_download",
    "vh_buffer_copy",
    "vh_grid_create", "vh_grid_destroy", "vh_grid_ncell", "vh_grid_set_mode", "vh_grid_get_mode", "vh_grid_upload", "vh_grid_download",
    "vh_grid_state_device_ptr", "vh_grid_compute_dt", "vh_grid_step", "vh_grid_run",
    "vh_grid_totals", "vh_grid_accumulate_flux", "vh_grid_sync", "vh_grid_set_launch",
    "vh_grid_tile_count", "vh_grid_get_launch", "vh_grid_time_steps",
    "vh_blocks_describe", "vh_blocks_create", "vh_blocks_destroy", "vh_blocks_ncell",
    "vh_blocks_set_mode", "vh_blocks_upload", "vh_blocks_download", "vh_blocks_compute_dt",
    "vh_blocks_step", "vh_blocks_run", "vh_blocks_totals", "vh_blocks_set_launch",
    "vh_blocks_info", "vh_blocks_sync", "vh_blocks_time_steps", "vh_blocks_set_tiling",
    "vh_blocks_indicator", "vh_blocks_remap", "vh_blocks_regrid", "vh_blocks_table_digest",
    "vh_blocks_enable_gravity", "vh_blocks_disable_gravity", "vh_blocks_solve_poisson", "vh_blocks_phi_upload",
    "vh_blocks_phi_download", "vh_blocks_apply_laplacian", "vh_blocks_gravity_accel",
    "vh_grid_step_vjp", "vh_adjoint_create", "vh_adjoint_destroy", "vh_adjoint_forward",
    "vh_adjoint_backward", "vh_adjoint_info", "vh_grid_set_adjoint_form",
    "vh_grid_step_bflux", "vh_grid_bflux_count", "vh_grid_set_parent_ghost",
    "vh_grid_enable_gravity", "vh_grid_disable_gravity", "vh_grid_solve_poisson", "vh_grid_phi_upload",
    "vh_grid_phi_download", "vh_grid_apply_laplacian", "vh_grid_gravity_accel",
    "vh_muscl_create", "vh_muscl_destroy", "vh_muscl_ncell", "vh_muscl_upload", "vh_muscl_download",
    "vh_muscl_state_device_ptr", "vh_muscl_set_tiles", "vh_muscl_step", "vh_muscl_fluxrec_download",
    "vh_muscl_fill_periodic", "vh_muscl_max_wavespeed", "vh_muscl_sync", "vh_muscl_time_steps",
    "vh_muscl_flux_register", "vh_muscl_enable_dual_energy", "vh_muscl_set_mode", "vh_muscl_set_bc",
)


class VespaHIPError(RuntimeError):
    """A C-ABI call returned a non-zero status (message from vh_last_error)."""


_lib = None


def load_library(path: str | None = None):
    """Load and type the HIP library. Raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(
            f"libvespahip.so not found at {p}: build it with vespa.jl_amd/build.py "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    L = C.CDLL(p)
    dp, fp, vp = C.POINTER(C.c_double), C.POINTER(C.c_float), C.c_void_p
    i32p, i64p = C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    L.vh_last_error.restype = C.c_char_p
    L.vh_abi_version.restype = C.c_int
    L.vh_device_count.restype = C.c_int
    L.vh_device_cu_count.argtypes = [C.POINTER(C.c_int32)]
    L.vh_device_cu_count.restype = C.c_int
    for name, tp in (("f64", dp), ("f32", fp)):
        f = getattr(L, f"vh_face_flux_{name}")
        f.argtypes = [tp, tp, i32p, C.c_int64, C.c_double, tp]
        f.restype = C.c_int
        f = getattr(L, f"vh_face_flux_dev_{name}")
        f.argtypes = [vp, vp, vp, C.c_int64, C.c_double, vp, vp]
        f.restype = C.c_int
    L.vh_face_flux_noalloc.argtypes = [vp, vp, i32p, C.c_int64, C.c_double, C.c_int32, vp, vp]
    L.vh_face_flux_noalloc.restype = C.c_int
    L.vh_buffer_alloc.argtypes = [C.POINTER(vp), C.c_int64]
    L.vh_buffer_free.argtypes = [vp]
    L.vh_buffer_upload.argtypes = [vp, vp, C.c_int64]
    L.vh_buffer_download.argtypes = [vp, vp, C.c_int64]
    L.vh_buffer_copy.argtypes = [vp, vp, C.c_int64]
    for n in ("vh_buffer_alloc", "vh_buffer_free", "vh_buffer_upload", "vh_buffer_download",
              "vh_buffer_copy"):
        getattr(L, n).restype = C.c_int
    L.vh_grid_create.argtypes = [C.POINTER(vp), C.c_int32, i64p, dp, dp, i32p, C.c_double, C.c_int32]
    L.vh_grid_destroy.argtypes = [vp]
    L.vh_grid_ncell.argtypes = [vp]
    L.vh_grid_ncell.restype = C.c_int64
    L.vh_grid_set_mode.argtypes = [vp, C.c_int32]
    L.vh_grid_set_mode.restype = C.c_int
    L.vh_grid_get_mode.argtypes = [vp]
    L.vh_grid_get_mode.restype = C.c_int
    L.vh_grid_upload.argtypes = [vp, vp]
    L.vh_grid_download.argtypes = [vp, vp]
    L.vh_grid_state_device_ptr.argtypes = [vp, C.POINTER(vp)]
    L.vh_grid_compute_dt.argtypes = [vp, C.c_double, dp]
    L.vh_grid_step.argtypes = [vp, C.c_double]
    L.vh_grid_run.argtypes = [vp, C.c_double, C.c_double, C.c_int64, dp, i64p]
    L.vh_grid_totals.argtypes = [vp, dp]
    L.vh_grid_accumulate_flux.argtypes = [vp, vp]
    L.vh_grid_sync.argtypes = [vp]
    L.vh_grid_set_launch.argtypes = [vp, C.c_int32, C.c_int32]
    L.vh_grid_tile_count.argtypes = [vp]
    L.vh_grid_get_launch.argtypes = [vp, i32p, i32p, i32p, i64p]
    L.vh_grid_time_steps.argtypes = [vp, C.c_double, C.c_int64, dp, dp, dp]
    for n in ("vh_grid_create", "vh_grid_destroy", "vh_grid_upload", "vh_grid_download",
              "vh_grid_state_device_ptr", "vh_grid_compute_dt", "vh_grid_step", "vh_grid_run",
              "vh_grid_totals", "vh_grid_accumulate_flux", "vh_grid_sync", "vh_grid_set_launch",
              "vh_grid_tile_count", "vh_grid_get_launch", "vh_grid_time_steps"):
        getattr(L, n).restype = C.c_int
    L.vh_blocks_describe.argtypes = [C.c_int32, C.c_int32, C.c_int32, dp, i32p, i32p, i64p, i32p, C.c_int32,
                                     i64p]
    L.vh_blocks_create.argtypes = [C.POINTER(vp), C.c_int32, C.c_int32, C.c_int32, dp, C.c_double, i32p, i32p,
                                   i64p, i32p, C.c_int32]
    L.vh_blocks_destroy.argtypes = [vp]
    L.vh_blocks_ncell.argtypes = [vp]
    L.vh_blocks_ncell.restype = C.c_int64
    L.vh_blocks_set_mode.argtypes = [vp, C.c_int32]
    L.vh_blocks_upload.argtypes = [vp, vp]
    L.vh_blocks_download.argtypes = [vp, vp]
    L.vh_blocks_compute_dt.argtypes = [vp, C.c_double, C.c_int32, dp]
    L.vh_blocks_step.argtypes = [vp, C.c_double, C.c_int32]
    L.vh_blocks_run.argtypes = [vp, C.c_double, C.c_double, C.c_int64, C.c_int32, dp, i64p]
    L.vh_blocks_totals.argtypes = [vp, dp]
    L.vh_blocks_set_launch.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32]
    L.vh_blocks_info.argtypes = [vp, i64p, i32p, i32p]
    L.vh_blocks_set_tiling.argtypes = [vp, C.c_int32]
    L.vh_blocks_sync.argtypes = [vp]
    L.vh_blocks_time_steps.argtypes = [vp, C.c_double, C.c_int64, C.c_int32, dp, dp]
    L.vh_blocks_indicator.argtypes = [vp, dp, dp]
    L.vh_blocks_remap.argtypes = [vp, vp, i32p]
    L.vh_blocks_regrid.argtypes = [vp, C.c_int32, i32p, i32p, i64p, i32p, i32p]
    L.vh_blocks_table_digest.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double), i32p, i32p, i64p,
                                         i32p, C.c_int32, C.c_int32, C.POINTER(C.c_uint64)]
    for n in ("vh_blocks_describe", "vh_blocks_create", "vh_blocks_destroy", "vh_blocks_set_mode",
              "vh_blocks_upload", "vh_blocks_download", "vh_blocks_compute_dt", "vh_blocks_step",
              "vh_blocks_run", "vh_blocks_totals", "vh_blocks_set_launch", "vh_blocks_info",
              "vh_blocks_sync", "vh_blocks_time_steps", "vh_blocks_set_tiling", "vh_blocks_indicator",
              "vh_blocks_remap", "vh_blocks_regrid", "vh_blocks_table_digest"):
        getattr(L, n).restype = C.c_int
    L.vh_grid_step_vjp.argtypes = [vp, vp, C.c_double, vp, vp]
    L.vh_adjoint_create.argtypes = [C.POINTER(vp), vp, C.c_int64, C.c_int32]
    L.vh_adjoint_destroy.argtypes = [vp]
    L.vh_adjoint_forward.argtypes = [vp, C.c_double]
    L.vh_adjoint_backward.argtypes = [vp, vp, vp]
    L.vh_adjoint_info.argtypes = [vp, dp, dp, i64p, dp]
    L.vh_grid_set_adjoint_form.argtypes = [vp, C.c_int32]
    L.vh_grid_step_bflux.argtypes = [vp, C.c_double, vp]
    L.vh_grid_step_bflux.restype = C.c_int
    L.vh_grid_set_parent_ghost.argtypes = [vp, vp]
    L.vh_grid_set_parent_ghost.restype = C.c_int
    L.vh_grid_bflux_count.argtypes = [vp]
    L.vh_grid_bflux_count.restype = C.c_int64
    L.vh_blocks_enable_gravity.argtypes = [vp, C.c_double, C.c_int64, C.c_double, C.c_int32]
    L.vh_blocks_disable_gravity.argtypes = [vp]
    L.vh_blocks_solve_poisson.argtypes = [vp, i64p, dp]
    L.vh_blocks_phi_upload.argtypes = [vp, vp]
    L.vh_blocks_phi_download.argtypes = [vp, vp]
    L.vh_blocks_apply_laplacian.argtypes = [vp, vp, vp]
    L.vh_blocks_gravity_accel.argtypes = [vp, vp]
    L.vh_grid_enable_gravity.argtypes = [vp, C.c_double, C.c_int32, C.c_int64, C.c_double, C.c_int32]
    L.vh_grid_disable_gravity.argtypes = [vp]
    L.vh_grid_solve_poisson.argtypes = [vp, i64p, dp]
    L.vh_grid_phi_upload.argtypes = [vp, vp]
    L.vh_grid_phi_download.argtypes = [vp, vp]
    L.vh_grid_apply_laplacian.argtypes = [vp, vp, vp]
    L.vh_grid_gravity_accel.argtypes = [vp, vp]
    L.vh_muscl_create.argtypes = [C.POINTER(vp), i32p, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double,
                                  C.c_int32]
    L.vh_muscl_destroy.argtypes = [vp]
    L.vh_muscl_ncell.argtypes = [vp]
    L.vh_muscl_ncell.restype = C.c_int64
    L.vh_muscl_upload.argtypes = [vp, vp]
    L.vh_muscl_download.argtypes = [vp, vp]
    L.vh_muscl_state_device_ptr.argtypes = [vp, C.POINTER(vp)]
    L.vh_muscl_set_tiles.argtypes = [vp, C.c_int32, C.c_int32]
    L.vh_muscl_step.argtypes = [vp, C.c_double, C.c_int32]
    L.vh_muscl_fluxrec_download.argtypes = [vp, vp]
    L.vh_muscl_fill_periodic.argtypes = [vp]
    L.vh_muscl_max_wavespeed.argtypes = [vp, dp]
    L.vh_muscl_sync.argtypes = [vp]
    L.vh_muscl_time_steps.argtypes = [vp, C.c_double, C.c_int64, C.c_int32, dp, dp]
    L.vh_muscl_flux_register.argtypes = [vp, C.c_double, dp, dp]
    L.vh_muscl_enable_dual_energy.argtypes = [vp, C.c_double]
    L.vh_muscl_set_mode.argtypes = [vp, C.c_int32]
    L.vh_muscl_set_bc.argtypes = [vp, C.c_int32, vp, vp]
    for n in ("vh_muscl_create", "vh_muscl_destroy", "vh_muscl_upload", "vh_muscl_download",
              "vh_muscl_state_device_ptr", "vh_muscl_set_tiles", "vh_muscl_step", "vh_muscl_fluxrec_download",
              "vh_muscl_fill_periodic", "vh_muscl_max_wavespeed", "vh_muscl_sync", "vh_muscl_time_steps",
              "vh_muscl_flux_register", "vh_muscl_enable_dual_energy", "vh_muscl_set_mode", "vh_muscl_set_bc"):
        getattr(L, n).restype = C.c_int
    for n in ("vh_grid_enable_gravity", "vh_grid_disable_gravity", "vh_grid_solve_poisson", "vh_grid_phi_upload",
              "vh_grid_phi_download", "vh_grid_apply_laplacian", "vh_grid_gravity_accel"):
        getattr(L, n).restype = C.c_int
    for n in ("vh_grid_step_vjp", "vh_adjoint_create", "vh_adjoint_destroy", "vh_adjoint_forward",
              "vh_adjoint_backward", "vh_adjoint_info", "vh_grid_set_adjoint_form"):
        getattr(L, n).restype = C.c_int
    if path is None:
        _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        msg = load_library().vh_last_error().decode(errors="replace")
        raise VespaHIPError(f"vespahip call failed ({rc}): {msg}")


class DeviceBuffer:
    """A raw device allocation owned by the library (hipMalloc/hipFree)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(load_library().vh_buffer_alloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    def upload(self, arr) -> "DeviceBuffer":
        import numpy as np
        a = np.ascontiguousarray(arr)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        check(load_library().vh_buffer_upload(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))
        return self

    def download(self, out):
        if out.nbytes > self.nbytes or not out.flags.c_contiguous:
            raise ValueError("output must be contiguous and fit the buffer")
        check(load_library().vh_buffer_download(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def copy_from(self, other: "DeviceBuffer", nbytes: int | None = None) -> "DeviceBuffer":
        n = min(self.nbytes, other.nbytes) if nbytes is None else int(nbytes)
        check(load_library().vh_buffer_copy(self.ptr, other.ptr, n))
        return self

    def free(self):
        if self.ptr:
            check(load_library().vh_buffer_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_count() -> int:
    return int(load_library().vh_device_count())


def device_cu_count() -> int:
    """Compute units of the current HIP device."""
    n = C.c_int32()
    check(load_library().vh_device_cu_count(C.byref(n)))
    return int(n.value)
