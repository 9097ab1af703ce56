"""Host mirror of PPMKernels' unsplit SSP-RK2 MUSCL driver
(lib/PPMKernels/src/muscl_grid.jl: muscl_step_3d!, prim_to_cons!, cons_to_prim!,
total_field, fill_periodic!, max_wavespeed) over the fused HIP stage kernels.

A :class:`MusclGrid` holds the conserved state (D, S1, S2, S3, τ = ρ·etot) of a
ghosted 3D grid on the device: ``dims`` counts cells per axis including ``ng``
ghost zones on each side, x fastest. ``step(dt)`` is ``muscl_step_3d!`` without a
``bc!`` callback (ghost zones are left as they are; call ``fill_periodic()``
between steps for a periodic box). With ``dual=True`` the state carries the gas
energy ρe as a sixth field (the ``ge``/``eta1`` dual-energy formalism).
``muscl_step_3d`` is the array-in/array-out form of the reference call.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, load_library


_BC_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32)


class MusclGrid:
    def __init__(self, dims, ng: int, gamma: float, theta: float = 1.5, dx: float = 1.0, small_rho: float = 1e-10,
                 eltype=np.float64, dual: bool = False, eta1: float = 1e-3):
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
        self.nf = 5
        self.mode = "exact"
        if dual:
            check(self._lib.vh_muscl_enable_dual_energy(self._h, float(eta1)))
            self.nf = 6

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
        if U.size != self.nf * self.ncell:
            raise ValueError(f"state must hold {self.nf}*{self.ncell} values")
        check(self._lib.vh_muscl_upload(self._h, U.ctypes.data_as(C.c_void_p)))

    def state(self) -> np.ndarray:
        out = np.empty((self.nf, self.ncell), dtype=self.eltype)
        check(self._lib.vh_muscl_download(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def set_mode(self, mode: str = "exact"):
        """Arithmetic profile of the stage kernels: "exact" (the reference's operations and
        order, bit-identical) or "fast" (contracted a*b+c, refined reciprocals and rsqrt
        sound speeds; within a few ulps per face)."""
        if mode not in ("exact", "fast"):
            raise ValueError(f"mode must be 'exact' or 'fast', got {mode!r}")
        check(self._lib.vh_muscl_set_mode(self._h, _lib.VH_EXACT if mode == "exact" else _lib.VH_FAST))
        self.mode = mode

    def set_tiles(self, variant: int = 0, zchunk: int = 8):
        """Tile shape: 0 = 16x16 cells, 1 = 32x8; zchunk z planes per workgroup."""
        check(self._lib.vh_muscl_set_tiles(self._h, int(variant), int(zchunk)))

    def step(self, dt: float, fluxrec: bool = False):
        """muscl_step_3d!(…; dt, fluxrec) (lib/PPMKernels/src/muscl_grid.jl:314-370)."""
        check(self._lib.vh_muscl_step(self._h, float(dt), int(bool(fluxrec))))

    def fluxrec(self) -> np.ndarray:
        """Fluxes recorded by the last recording step: (3, nf, ncell), [axis][field][cell],
        ½(F₁+F₂) through the lo face of each cell (grid-order momenta)."""
        out = np.empty((3, self.nf, self.ncell), dtype=self.eltype)
        check(self._lib.vh_muscl_fluxrec_download(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def fill_periodic(self):
        """fill_periodic!(dims, ng, D, S1, S2, S3, Tau) (lib/PPMKernels/src/muscl_grid.jl:664-670)."""
        check(self._lib.vh_muscl_fill_periodic(self._h))

    def set_bc(self, bc=None):
        """bc! of muscl_step_3d! (lib/PPMKernels/src/muscl_grid.jl:316-322, 339, 350),
        run before each RK stage: None (bc! = nothing), "periodic" (fill_periodic!
        on the device), or a host function bc(V, stage) that refills the ghost
        zones of V, the stage's input as an (nf, nz, ny, nx) array, in place (the
        state crosses PCIe twice per stage; a device-side fill belongs in the C
        ABI's callback, which receives the device pointer and the stream)."""
        if bc is None:
            check(self._lib.vh_muscl_set_bc(self._h, 0, None, None))
            self._bc_keep = None
        elif isinstance(bc, str):
            if bc != "periodic":
                raise ValueError(f"unknown bc {bc!r}")
            check(self._lib.vh_muscl_set_bc(self._h, 1, None, None))
            self._bc_keep = None
        else:
            nf, shape = self.nf, (self.nf,) + self.dims[::-1]
            nbytes = nf * self.ncell * self.eltype.itemsize
            lib, h, dt = self._lib, self._h, self.eltype

            def hook(_user, state, _stream, stage):
                try:
                    lib.vh_muscl_sync(h)
                    host = np.empty(shape, dtype=dt)
                    if lib.vh_buffer_download(host.ctypes.data_as(C.c_void_p), state, nbytes):
                        return 2
                    bc(host, int(stage))
                    host = np.ascontiguousarray(host, dtype=dt)
                    return 3 if lib.vh_buffer_upload(state, host.ctypes.data_as(C.c_void_p), nbytes) else 0
                except Exception:   # the error surfaces as the step's failure
                    return 1

            cb = _BC_FN(hook)
            self._bc_keep = cb
            check(self._lib.vh_muscl_set_bc(self._h, 2, C.cast(cb, C.c_void_p), None))

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


class BoundaryFluxSet:
    """The reflux fluxes of one grid (lib/PPMKernels/src/reflux_capture.jl:38-50):
    ``flux[(axis, side, cell)]`` on the outer boundary faces, side "lo"/"hi", and
    ``interior[(axis, cell)]`` on the hi face of each active cell but the last along
    the axis; values are 5-tuples of ∫F·area·dt. Axes and cell indices are 1-based
    active indices (the reference's keys); ``act`` the active dims."""

    def __init__(self, flux: dict, interior: dict, act: tuple):
        self.flux, self.interior, self.act = flux, interior, act


def _trans3(a):
    return (2, 3) if a == 1 else ((1, 3) if a == 2 else (1, 2))


def _key(a, k, b1, b2, u1, u2):
    return tuple(k if d == a else (u1 if d == b1 else u2) for d in (1, 2, 3))


def flux_register_arrays(g: MusclGrid, dt: float, interior: bool = True):
    """vh_muscl_flux_register: per axis (planes (2, 5, act_b2, act_b1), interior
    (5, act_b2, act_b1, act_a − 1) or None), Float64 ∫F·area·dt."""
    act = [d - 2 * g.ng for d in g.dims]
    sizes = []
    for a in range(3):
        b1, b2 = [d for d in range(3) if d != a]
        sizes.append((act[b1], act[b2], act[a]))
    npl = sum(10 * m1 * m2 for m1, m2, _ in sizes)
    nin = sum(5 * (ma - 1) * m1 * m2 for m1, m2, ma in sizes)
    pl = np.empty(npl)
    inn = np.empty(max(nin, 1)) if interior else None
    dp = C.POINTER(C.c_double)
    check(g._lib.vh_muscl_flux_register(g._h, float(dt), pl.ctypes.data_as(dp),
                                        inn.ctypes.data_as(dp) if interior else None))
    out, op, oi = [], 0, 0
    for m1, m2, ma in sizes:
        p = pl[op:op + 10 * m1 * m2].reshape(2, 5, m2, m1)
        op += 10 * m1 * m2
        i = None
        if interior:
            i = inn[oi:oi + 5 * (ma - 1) * m1 * m2].reshape(5, m2, m1, ma - 1)
            oi += 5 * (ma - 1) * m1 * m2
        out.append((p, i))
    return out


def boundary_flux_register(g: MusclGrid, dt: float) -> BoundaryFluxSet:
    """boundary_flux_register(frec, gd, ng, dt, dx) (lib/PPMKernels/src/reflux_capture.jl:86-131)
    for the fluxes recorded by the grid's last ``step(dt, fluxrec=True)``: the device
    gathers the face planes, the host keys them."""
    act = tuple(d - 2 * g.ng for d in g.dims)
    flux, interior = {}, {}
    for a, (p, inn) in enumerate(flux_register_arrays(g, dt), start=1):
        b1, b2 = _trans3(a)
        ma, m1, m2 = act[a - 1], act[b1 - 1], act[b2 - 1]
        for u2 in range(m2):
            for u1 in range(m1):
                flux[(a, "lo", _key(a, 1, b1, b2, u1 + 1, u2 + 1))] = tuple(float(v) for v in p[0, :, u2, u1])
                flux[(a, "hi", _key(a, ma, b1, b2, u1 + 1, u2 + 1))] = tuple(float(v) for v in p[1, :, u2, u1])
                for c in range(ma - 1):
                    interior[(a, _key(a, c + 1, b1, b2, u1 + 1, u2 + 1))] = tuple(float(v) for v in inn[:, u2, u1, c])
    return BoundaryFluxSet(flux, interior, act)


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
                  bc=None, ge=None, eta1=1e-3, fluxrec=None):
    """muscl_step_3d!(D, S1, S2, S3, Tau, dims, ng; dt, gamma, theta, dx, small_rho, bc!,
    ge, eta1, fluxrec) on host arrays, updated in place through a device context.
    bc: None, "periodic", or bc!(D, S1, S2, S3, Tau[, ge]) on flat host field
    arrays, called before each stage. fluxrec: None or a (3, nf, N) array filled
    with the recorded fluxes (nf = 6 with ge)."""
    fields = (D, S1, S2, S3, Tau) + ((ge,) if ge is not None else ())
    g = MusclGrid(dims, ng, gamma, theta, dx, small_rho, eltype=np.asarray(D).dtype, dual=ge is not None,
                  eta1=eta1)
    try:
        if callable(bc):
            g.set_bc(lambda V, stage: bc(*[f.reshape(-1) for f in V]))
        else:
            g.set_bc(bc)
        g.upload(np.stack([np.ravel(f) for f in fields]))
        g.step(dt, fluxrec is not None)
        U = g.state()
        for f, u in zip(fields, U):
            np.copyto(f, u.reshape(np.shape(f)))
        if fluxrec is not None:
            fluxrec[...] = g.fluxrec()
    finally:
        g.close()
