"""Host mirror of PPMKernels' unsplit SSP-RK2 MUSCL driver
(lib/PPMKernels/src/muscl_grid.jl: muscl_step_3d!, prim_to_cons!, cons_to_prim!,
total_field, fill_periodic!, max_wavespeed) over the fused HIP stage kernels.

A :class:`MusclGrid` holds the conserved state (D, S1, S2, S3, τ = ρ·etot) of a
ghosted 3D grid on the device: ``dims`` counts cells per axis including ``ng``
ghost zones on each side, x fastest. ``step(dt)`` is ``muscl_step_3d!`` without a
``bc!`` callback (ghost zones are left as they are; call ``fill_periodic()``
between steps for a periodic box) and without the dual-energy field.
``muscl_step_3d`` is the array-in/array-out form of the reference call.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, load_library


class MusclGrid:
    def __init__(self, dims, ng: int, gamma: float, theta: float = 1.5, dx: float = 1.0, small_rho: float = 1e-10,
                 eltype=np.float64):
        self.dims = tuple(int(d) for d in dims)
        if len(self.dims) != 3:
            raise ValueError("muscl_step_3d! works on 3D grids")
        self.ng = int(ng)
        self.eltype = np.dtype(eltype)
        if self.eltype not in (np.float64, np.float32):
            raise ValueError(f"eltype must be float64 or float32, got {self.eltype}")
        self.gamma, self.theta, self.dx, self.small_rho = float(gamma), float(theta), float(dx), float(small_rho)
        self._lib = load_library()
        h = C.c_void_p()
        d = (C.c_int32 * 3)(*self.dims)
        check(self._lib.vh_muscl_create(C.byref(h), d, self.ng, self.gamma, self.theta, self.small_rho, self.dx,
                                        _lib.VH_F64 if self.eltype == np.float64 else _lib.VH_F32))
        self._h = h
        self.ncell = int(np.prod(self.dims))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.vh_muscl_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, U: np.ndarray):
        U = np.ascontiguousarray(U, dtype=self.eltype)
        if U.size != 5 * self.ncell:
            raise ValueError(f"state must hold 5*{self.ncell} values")
        check(self._lib.vh_muscl_upload(self._h, U.ctypes.data_as(C.c_void_p)))

    def state(self) -> np.ndarray:
        out = np.empty((5, self.ncell), dtype=self.eltype)
        check(self._lib.vh_muscl_download(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def set_tiles(self, variant: int = 0, zchunk: int = 8):
        """Tile shape: 0 = 16x16 cells, 1 = 32x8; zchunk z planes per workgroup."""
        check(self._lib.vh_muscl_set_tiles(self._h, int(variant), int(zchunk)))

    def step(self, dt: float, fluxrec: bool = False):
        """muscl_step_3d!(…; dt, fluxrec) (lib/PPMKernels/src/muscl_grid.jl:314-370)."""
        check(self._lib.vh_muscl_step(self._h, float(dt), int(bool(fluxrec))))

    def fluxrec(self) -> np.ndarray:
        """Fluxes recorded by the last recording step: (3, 5, ncell), [axis][field][cell],
        ½(F₁+F₂) through the lo face of each cell (grid-order momenta)."""
        out = np.empty((3, 5, self.ncell), dtype=self.eltype)
        check(self._lib.vh_muscl_fluxrec_download(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def fill_periodic(self):
        """fill_periodic!(dims, ng, D, S1, S2, S3, Tau) (lib/PPMKernels/src/muscl_grid.jl:664-670)."""
        check(self._lib.vh_muscl_fill_periodic(self._h))

    def max_wavespeed(self) -> float:
        """max_wavespeed (lib/PPMKernels/src/muscl_grid.jl:689-695)."""
        out = C.c_double()
        check(self._lib.vh_muscl_max_wavespeed(self._h, C.byref(out)))
        return out.value

    def sync(self):
        check(self._lib.vh_muscl_sync(self._h))

    def time_steps(self, dt: float, nsteps: int, fluxrec: bool = False) -> dict:
        a, b = C.c_double(), C.c_double()
        check(self._lib.vh_muscl_time_steps(self._h, float(dt), int(nsteps), int(bool(fluxrec)), C.byref(a),
                                            C.byref(b)))
        return {"step_ms": a.value, "stage2_ms": b.value}


def prim_to_cons(rho, vx, vy, vz, etot):
    """prim_to_cons! (lib/PPMKernels/src/muscl_grid.jl:612-619) → (5, N)."""
    return np.stack([rho, rho * vx, rho * vy, rho * vz, rho * etot])


def cons_to_prim(U):
    """cons_to_prim! (lib/PPMKernels/src/muscl_grid.jl:622-629) → (rho, vx, vy, vz, eint)."""
    D = U[0]
    half = U.dtype.type(0.5)
    return (D, U[1] / D, U[2] / D, U[3] / D,
            U[4] / D - half * ((U[1] / D) ** 2 + (U[2] / D) ** 2 + (U[3] / D) ** 2))


def total_field(f, dims, ng, dx):
    """total_field (lib/PPMKernels/src/muscl_grid.jl:632-639): Σ f·dV over the active cells."""
    nx, ny, nz = dims
    a = np.asarray(f).reshape(nz, ny, nx)[ng:nz - ng, ng:ny - ng, ng:nx - ng]
    return a.sum() * (type(a.flat[0])(dx) ** 3)


def muscl_step_3d(D, S1, S2, S3, Tau, dims, ng, *, dt, gamma, theta=1.5, dx=1.0, small_rho=1e-10,
                  fluxrec=None):
    """muscl_step_3d!(D, S1, S2, S3, Tau, dims, ng; dt, gamma, theta, dx, small_rho,
    fluxrec) on host arrays, updated in place through a device context. fluxrec: None
    or a (3, 5, N) array filled with the recorded fluxes."""
    fields = (D, S1, S2, S3, Tau)
    g = MusclGrid(dims, ng, gamma, theta, dx, small_rho, eltype=np.asarray(D).dtype)
    try:
        g.upload(np.stack([np.ravel(f) for f in fields]))
        g.step(dt, fluxrec is not None)
        U = g.state()
        for f, u in zip(fields, U):
            np.copyto(f, u.reshape(np.shape(f)))
        if fluxrec is not None:
            fluxrec[...] = g.fluxrec()
    finally:
        g.close()
