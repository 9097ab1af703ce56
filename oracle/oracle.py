"""ctypes wrapper over the CPU oracle ``liboracle.so`` (TEST INFRASTRUCTURE ONLY).

The oracle is the checker: only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module. The product path
(``vespa.jl_amd``) never imports it.

Arrays follow the reference's SoA storage (lib/RefMesh/src/RefMesh.jl:220-226):
one contiguous block of 5 fields, each field column-major (x fastest), so a numpy
array of shape ``(5, nz, ny, nx)`` in C order is exactly that layout.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

BC_PERIODIC, BC_OUTFLOW, BC_REFLECTING, BC_PARENT = 0, 1, 2, 3
_BC_NAMES = {"periodic": BC_PERIODIC, "outflow": BC_OUTFLOW, "reflecting": BC_REFLECTING, "parent": BC_PARENT}


class Grid(C.Structure):
    _fields_ = [
        ("rank", C.c_int32),
        ("bc", C.c_int32 * 6),
        ("dims", C.c_int64 * 3),
        ("lo", C.c_double * 3),
        ("hi", C.c_double * 3),
        ("gamma", C.c_double),
        ("pg", C.c_void_p),
    ]


def build() -> None:
    """Compile liboracle.so from oracle/vespa_oracle.c and oracle/jvp_oracle.cc
    (gcc/g++, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-B", "-C", HERE, "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp, fp = C.POINTER(C.c_double), C.POINTER(C.c_float)
        gp = C.POINTER(Grid)
        for suf, tp in (("f64", dp), ("f32", fp)):
            getattr(L, f"vo_hllc_flux_{suf}").argtypes = [tp, tp, C.c_double, C.c_int, tp]
            getattr(L, f"vo_euler_flux_{suf}").argtypes = [tp, C.c_double, C.c_int, tp]
            getattr(L, f"vo_cons2prim_{suf}").argtypes = [tp, C.c_double, C.c_int, tp]
            getattr(L, f"vo_prim2cons_{suf}").argtypes = [tp, C.c_double, tp]
            getattr(L, f"vo_face_flux_{suf}").argtypes = [tp, tp, C.POINTER(C.c_int32), C.c_int64,
                                                          C.c_double, tp]
            getattr(L, f"vo_accumulate_flux_{suf}").argtypes = [gp, tp, tp, C.c_int]
            getattr(L, f"vo_accumulate_flux_{suf}").restype = C.c_int
            getattr(L, f"vo_step_{suf}").argtypes = [gp, tp, C.c_double, tp, C.c_int, tp]
            getattr(L, f"vo_step_{suf}").restype = C.c_int
            getattr(L, f"vo_compute_dt_{suf}").argtypes = [gp, tp, C.c_double, C.c_int]
            getattr(L, f"vo_compute_dt_{suf}").restype = C.c_double
            getattr(L, f"vo_totals_{suf}").argtypes = [gp, tp, dp]
            getattr(L, f"vo_evolve_{suf}").argtypes = [gp, tp, C.c_double, C.c_double, C.c_int64,
                                                       C.c_int, dp]
            getattr(L, f"vo_evolve_{suf}").restype = C.c_int64
            getattr(L, f"vo_run_steps_{suf}").argtypes = [gp, tp, C.c_int64, C.c_double, C.c_int, dp]
            getattr(L, f"vo_run_steps_{suf}").restype = C.c_int
            getattr(L, f"vo_accumulate_flux_gather_{suf}").argtypes = [gp, tp, tp, C.c_int]
            getattr(L, f"vo_accumulate_flux_gather_{suf}").restype = C.c_int
            getattr(L, f"vo_step_omp_{suf}").argtypes = [gp, tp, C.c_double, tp, C.c_int, C.c_int]
            getattr(L, f"vo_step_omp_{suf}").restype = C.c_int
            getattr(L, f"vo_compute_dt_omp_{suf}").argtypes = [gp, tp, C.c_double, C.c_int, C.c_int]
            getattr(L, f"vo_compute_dt_omp_{suf}").restype = C.c_double
            getattr(L, f"vo_poisson_apply_{suf}").argtypes = [gp, C.c_int, tp, tp]
            getattr(L, f"vo_poisson_rhs_{suf}").argtypes = [gp, tp, C.c_double, C.c_int, tp]
            getattr(L, f"vo_poisson_solve_{suf}").argtypes = [gp, C.c_int, tp, C.c_double, C.c_int, C.c_int64,
                                                              C.c_double, tp, tp, dp]
            getattr(L, f"vo_poisson_solve_{suf}").restype = C.c_int64
            getattr(L, f"vo_grav_accel_{suf}").argtypes = [gp, C.c_int, tp, tp]
            getattr(L, f"vo_step_grav_{suf}").argtypes = [gp, tp, C.c_double, tp, C.c_int, C.c_int, tp]
            getattr(L, f"vo_step_grav_{suf}").restype = C.c_int
            getattr(L, f"vo_compute_dt_grav_{suf}").argtypes = [gp, tp, C.c_double, C.c_int, C.c_double]
            getattr(L, f"vo_compute_dt_grav_{suf}").restype = C.c_double
        L.vo_minmod_f64.argtypes = [C.c_double, C.c_double]
        L.vo_minmod_f64.restype = C.c_double
        L.vo_minmod_f32.argtypes = [C.c_float, C.c_float]
        L.vo_minmod_f32.restype = C.c_float
        L.vo_bface_count.argtypes = [gp]
        L.vo_bface_count.restype = C.c_int64
        L.vo_star_state.argtypes = [dp, dp, C.c_double, dp, dp]
        L.vo_accumulate_vjp_f64.argtypes = [gp, dp, dp, dp]
        L.vo_accumulate_vjp_f64.restype = C.c_int
        L.vo_step_vjp_f64.argtypes = [gp, dp, C.c_double, dp, dp]
        L.vo_step_vjp_f64.restype = C.c_int
        L.vo_c2p_vjp_f64.argtypes = [dp, C.c_double, dp, dp]
        L.vo_hllc_vjp_f64.argtypes = [dp, dp, C.c_double, C.c_int, dp, dp, dp]
        L.vo_exact_riemann_sample.argtypes = [dp, dp, C.c_double, C.c_double, dp]
        L.vo_jvp_max_directions.restype = C.c_int
        L.vo_rollout_jvp_f64.argtypes = [gp, dp, C.c_int64, dp, C.c_int, dp, C.c_int, dp, dp]
        L.vo_rollout_jvp_f64.restype = C.c_int
        L.vo_c2p_jvp_f64.argtypes = [dp, C.c_double, dp, dp, dp]
        L.vo_hllc_jvp_f64.argtypes = [dp, dp, C.c_double, C.c_int, dp, dp, dp, dp]
        _lib = L
    return _lib


def _suf(dtype) -> str:
    dtype = np.dtype(dtype)
    if dtype == np.float64:
        return "f64"
    if dtype == np.float32:
        return "f32"
    raise TypeError(f"unsupported dtype {dtype}")


def _ptr(a: np.ndarray):
    assert a.flags.c_contiguous
    tp = C.c_double if a.dtype == np.float64 else C.c_float
    return a.ctypes.data_as(C.POINTER(tp))


def make_grid(dims, domain, bcs, gamma) -> Grid:
    """dims: tuple of 1..3 ints (x first); domain: ((lo, hi), ...); bcs: one name
    for every side, or a list of (lo, hi) names per axis."""
    g = Grid()
    rank = len(dims)
    g.rank = rank
    for d in range(3):
        g.dims[d] = int(dims[d]) if d < rank else 1
        g.lo[d] = float(domain[d][0]) if d < rank else 0.0
        g.hi[d] = float(domain[d][1]) if d < rank else 1.0
    if isinstance(bcs, str):
        bcs = [(bcs, bcs)] * rank
    for d in range(3):
        lo, hi = bcs[d] if d < rank else ("outflow", "outflow")
        g.bc[2 * d] = _BC_NAMES[lo] if isinstance(lo, str) else int(lo)
        g.bc[2 * d + 1] = _BC_NAMES[hi] if isinstance(hi, str) else int(hi)
    g.gamma = float(gamma)
    g.pg = None
    return g


def set_parent_ghost(g: Grid, pg: np.ndarray) -> np.ndarray:
    """Attach ParentGhost states to g: pg is (nbface, 5) primitive states at
    the state's precision in the boundary-face slots (axis 0 lo, axis 0 hi,
    axis 1 lo, ...; each plane column-major over the other axes). Returns the
    array kept alive for the grid's use."""
    pg = np.ascontiguousarray(pg)
    g.pg = pg.ctypes.data
    g._pg_keep = pg
    return pg


def bface_planes(g: Grid):
    """(axis, side, first slot, plane shape (other axes, fastest first)) of every
    boundary plane in slot order."""
    out, base = [], 0
    dims = [int(g.dims[d]) for d in range(3)]
    for d in range(int(g.rank)):
        other = [dims[e] for e in range(3) if e != d]
        size = other[0] * other[1]
        for side in range(2):
            out.append((d, side, base, tuple(other)))
            base += size
    return out


def ncell(g: Grid) -> int:
    return int(g.dims[0] * g.dims[1] * g.dims[2])


# ---------------- scalar kernels ----------------

def hllc_flux(WL, WR, gamma, axis, dtype=np.float64):
    WL = np.ascontiguousarray(WL, dtype=dtype)
    WR = np.ascontiguousarray(WR, dtype=dtype)
    F = np.zeros(5, dtype=dtype)
    getattr(lib(), f"vo_hllc_flux_{_suf(dtype)}")(_ptr(WL), _ptr(WR), gamma, int(axis), _ptr(F))
    return F


def euler_flux(W, gamma, axis, dtype=np.float64):
    W = np.ascontiguousarray(W, dtype=dtype)
    F = np.zeros(5, dtype=dtype)
    getattr(lib(), f"vo_euler_flux_{_suf(dtype)}")(_ptr(W), gamma, int(axis), _ptr(F))
    return F


def cons2prim(U, gamma, dtype=np.float64, floor64=False):
    U = np.ascontiguousarray(U, dtype=dtype)
    W = np.zeros(5, dtype=dtype)
    getattr(lib(), f"vo_cons2prim_{_suf(dtype)}")(_ptr(U), gamma, int(floor64), _ptr(W))
    return W


def prim2cons(W, gamma, dtype=np.float64):
    W = np.ascontiguousarray(W, dtype=dtype)
    U = np.zeros(5, dtype=dtype)
    getattr(lib(), f"vo_prim2cons_{_suf(dtype)}")(_ptr(W), gamma, _ptr(U))
    return U


def minmod(a, b, dtype=np.float64):
    f = lib().vo_minmod_f64 if np.dtype(dtype) == np.float64 else lib().vo_minmod_f32
    return f(a, b)


def face_flux(WL, WR, axis, gamma):
    """Batched HLLC (src/ka_flux.jl:48-53): WL, WR (nf, 5); axis (nf,) in 1..3."""
    WL = np.ascontiguousarray(WL)
    WR = np.ascontiguousarray(WR, dtype=WL.dtype)
    ax = np.ascontiguousarray(axis, dtype=np.int32)
    nf = WL.shape[0]
    F = np.zeros_like(WL)
    getattr(lib(), f"vo_face_flux_{_suf(WL.dtype)}")(
        _ptr(WL), _ptr(WR), ax.ctypes.data_as(C.POINTER(C.c_int32)), nf, gamma, _ptr(F))
    return F


# ---------------- uniform-grid path ----------------

def accumulate_flux(g: Grid, U: np.ndarray, floor64=False) -> np.ndarray:
    U = np.ascontiguousarray(U)
    acc = np.zeros_like(U)
    rc = getattr(lib(), f"vo_accumulate_flux_{_suf(U.dtype)}")(C.byref(g), _ptr(U), _ptr(acc),
                                                               int(floor64))
    if rc != 0:
        raise ValueError("invalid grid")
    return acc


def step(g: Grid, U: np.ndarray, dt: float, floor64=False, bflux=None) -> None:
    """In-place SSP-RK2 step of U (src/driver.jl:333-359)."""
    assert U.flags.c_contiguous
    work = np.empty(10 * ncell(g), dtype=U.dtype)
    bp = _ptr(bflux) if bflux is not None else None
    rc = getattr(lib(), f"vo_step_{_suf(U.dtype)}")(C.byref(g), _ptr(U), float(dt), _ptr(work),
                                                   int(floor64), bp)
    if rc != 0:
        raise ValueError("invalid grid")


# ---------------- reverse mode (Float64, dt fixed) ----------------

def accumulate_vjp(g: Grid, U: np.ndarray, lam: np.ndarray) -> np.ndarray:
    """A'(U)^T lam for A = accumulate_flux."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    out = np.zeros_like(U)
    if lib().vo_accumulate_vjp_f64(C.byref(g), _ptr(U), _ptr(lam), _ptr(out)) != 0:
        raise ValueError("invalid grid")
    return out


def step_vjp(g: Grid, U: np.ndarray, dt: float, wbar: np.ndarray) -> np.ndarray:
    """dL/dU for L depending on step(U, dt) through wbar = dL/d(step(U, dt))."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    wbar = np.ascontiguousarray(wbar, dtype=np.float64)
    out = np.zeros_like(U)
    if lib().vo_step_vjp_f64(C.byref(g), _ptr(U), float(dt), _ptr(wbar), _ptr(out)) != 0:
        raise ValueError("invalid grid")
    return out


def rollout_jvp(g: Grid, U: np.ndarray, dts, V: np.ndarray, threads: int = 0):
    """Forward mode of len(dts) steps with the given dts held fixed: returns
    (final state, J·V) for tangent directions V shaped (ndir, *U.shape), ndir <=
    jvp_max_directions(). The final state equals the oracle's steps bit for bit;
    non-smooth points take the reverse mode's branches (jvp_oracle.cc)."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    if V.ndim != U.ndim + 1 or V.shape[1:] != U.shape:
        raise ValueError("V must be (ndir, *U.shape)")
    dts = np.ascontiguousarray(np.atleast_1d(dts), dtype=np.float64)
    ndir = V.shape[0]
    out = np.empty_like(U)
    JV = np.empty_like(V)
    if lib().vo_rollout_jvp_f64(C.byref(g), _ptr(U), len(dts), _ptr(dts), ndir, _ptr(V), int(threads),
                                _ptr(out), _ptr(JV)) != 0:
        raise ValueError("invalid grid or direction count")
    return out, JV


def jvp_max_directions() -> int:
    return int(lib().vo_jvp_max_directions())


def cons2prim_jvp(U, gamma, dU):
    U, dU = (np.ascontiguousarray(a, dtype=np.float64) for a in (U, dU))
    W, dW = np.zeros(5), np.zeros(5)
    lib().vo_c2p_jvp_f64(_ptr(U), float(gamma), _ptr(dU), _ptr(W), _ptr(dW))
    return W, dW


def hllc_jvp(WL, WR, gamma, axis, dWL, dWR):
    WL, WR, dWL, dWR = (np.ascontiguousarray(a, dtype=np.float64) for a in (WL, WR, dWL, dWR))
    F, dF = np.zeros(5), np.zeros(5)
    lib().vo_hllc_jvp_f64(_ptr(WL), _ptr(WR), float(gamma), int(axis), _ptr(dWL), _ptr(dWR), _ptr(F), _ptr(dF))
    return F, dF


def cons2prim_vjp(U, gamma, Wbar):
    U = np.ascontiguousarray(U, dtype=np.float64)
    Wbar = np.ascontiguousarray(Wbar, dtype=np.float64)
    out = np.zeros(5)
    lib().vo_c2p_vjp_f64(_ptr(U), float(gamma), _ptr(Wbar), _ptr(out))
    return out


def hllc_vjp(WL, WR, gamma, axis, Fbar):
    WL, WR, Fbar = (np.ascontiguousarray(a, dtype=np.float64) for a in (WL, WR, Fbar))
    a, b = np.zeros(5), np.zeros(5)
    lib().vo_hllc_vjp_f64(_ptr(WL), _ptr(WR), float(gamma), int(axis), _ptr(Fbar), _ptr(a), _ptr(b))
    return a, b


def accumulate_flux_gather(g: Grid, U: np.ndarray, floor64=False) -> np.ndarray:
    U = np.ascontiguousarray(U)
    acc = np.zeros_like(U)
    rc = getattr(lib(), f"vo_accumulate_flux_gather_{_suf(U.dtype)}")(C.byref(g), _ptr(U), _ptr(acc),
                                                                      int(floor64))
    if rc != 0:
        raise ValueError("invalid grid")
    return acc


def step_omp(g: Grid, U: np.ndarray, dt: float, floor64=False, threads=0, work=None) -> None:
    """In-place SSP-RK2 step in gather form on `threads` OpenMP threads."""
    assert U.flags.c_contiguous
    if work is None:
        work = np.empty(10 * ncell(g), dtype=U.dtype)
    rc = getattr(lib(), f"vo_step_omp_{_suf(U.dtype)}")(C.byref(g), _ptr(U), float(dt), _ptr(work),
                                                       int(floor64), int(threads))
    if rc != 0:
        raise ValueError("invalid grid")


def compute_dt_omp(g: Grid, U: np.ndarray, cfl: float, floor64=False, threads=0) -> float:
    U = np.ascontiguousarray(U)
    return getattr(lib(), f"vo_compute_dt_omp_{_suf(U.dtype)}")(C.byref(g), _ptr(U), cfl, int(floor64),
                                                               int(threads))


def compute_dt(g: Grid, U: np.ndarray, cfl: float, floor64=False) -> float:
    U = np.ascontiguousarray(U)
    return getattr(lib(), f"vo_compute_dt_{_suf(U.dtype)}")(C.byref(g), _ptr(U), cfl, int(floor64))


def totals(g: Grid, U: np.ndarray) -> np.ndarray:
    U = np.ascontiguousarray(U)
    out = np.zeros(5)
    getattr(lib(), f"vo_totals_{_suf(U.dtype)}")(C.byref(g), _ptr(U),
                                                out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def evolve(g: Grid, U: np.ndarray, tfinal: float, cfl: float, max_steps=0, floor64=False, t0=0.0):
    """In-place evolve! (src/driver.jl:483-508); returns (t, steps)."""
    assert U.flags.c_contiguous
    t = C.c_double(t0)
    k = getattr(lib(), f"vo_evolve_{_suf(U.dtype)}")(C.byref(g), _ptr(U), tfinal, cfl,
                                                    int(max_steps), int(floor64), C.byref(t))
    if k < 0:
        raise ValueError("invalid grid")
    return t.value, int(k)


def run_steps(g: Grid, U: np.ndarray, nsteps: int, cfl: float, floor64=False):
    """In-place: nsteps steps, dt from the CFL condition before each. Returns dts."""
    assert U.flags.c_contiguous
    dts = np.zeros(max(nsteps, 1))
    rc = getattr(lib(), f"vo_run_steps_{_suf(U.dtype)}")(
        C.byref(g), _ptr(U), int(nsteps), cfl, int(floor64), dts.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise ValueError("invalid grid")
    return dts[:nsteps]


def bface_count(g: Grid) -> int:
    return int(lib().vo_bface_count(C.byref(g)))


# ---------------- exact Riemann ----------------

def star_state(WL, WR, gamma):
    a = np.ascontiguousarray(WL, dtype=np.float64)
    b = np.ascontiguousarray(WR, dtype=np.float64)
    p, u = C.c_double(), C.c_double()
    lib().vo_star_state(_ptr(a), _ptr(b), gamma, C.byref(p), C.byref(u))
    return p.value, u.value


def exact_riemann_sample(WL, WR, gamma, s):
    a = np.ascontiguousarray(WL, dtype=np.float64)
    b = np.ascontiguousarray(WR, dtype=np.float64)
    out = np.zeros(3)
    lib().vo_exact_riemann_sample(_ptr(a), _ptr(b), gamma, float(s), _ptr(out))
    return out


# ---------------- block-structured AMR (blocks_oracle.c) ----------------

class BlockOracle:
    """Block-mesh restatement: `mesh` is a vespa.jl_amd.blockmesh.BlockMesh (or
    anything with the same fields)."""

    def __init__(self, mesh):
        L = lib()
        if not getattr(L, "_bo_typed", False):
            vp = C.c_void_p
            L.bo_create.restype = vp
            L.bo_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.c_double,
                                    vp, vp, vp, vp]
            L.bo_destroy.argtypes = [vp]
            L.bo_ncell.argtypes = [vp]
            L.bo_ncell.restype = C.c_int64
            L.bo_validate.argtypes = [vp]
            L.bo_validate.restype = C.c_int
            for suf in ("f64", "f32"):
                getattr(L, f"bo_accumulate_flux_{suf}").argtypes = [vp, vp, vp, C.c_int]
                getattr(L, f"bo_step_{suf}").argtypes = [vp, vp, C.c_double, C.c_int, vp, C.c_int]
                getattr(L, f"bo_compute_dt_{suf}").argtypes = [vp, vp, C.c_double, C.c_int, C.c_int]
                getattr(L, f"bo_compute_dt_{suf}").restype = C.c_double
                getattr(L, f"bo_totals_{suf}").argtypes = [vp, vp, vp]
                getattr(L, f"bo_evolve_{suf}").argtypes = [vp, vp, C.c_double, C.c_double, C.c_int64,
                                                           C.c_int, C.c_int, C.POINTER(C.c_double)]
                getattr(L, f"bo_evolve_{suf}").restype = C.c_int64
                getattr(L, f"bo_indicator_{suf}").argtypes = [vp, vp, vp, vp]
            L.bo_poisson_create.restype = vp
            L.bo_poisson_create.argtypes = [vp]
            L.bo_poisson_destroy.argtypes = [vp]
            L.bo_poisson_nnz.argtypes = [vp]
            L.bo_poisson_nnz.restype = C.c_int64
            L.bo_poisson_export.argtypes = [vp] * 7
            for suf in ("f64", "f32"):
                getattr(L, f"bo_poisson_apply_{suf}").argtypes = [vp, vp, vp]
                getattr(L, f"bo_poisson_rhs_{suf}").argtypes = [vp, vp, C.c_double, C.c_int, vp]
                getattr(L, f"bo_poisson_solve_{suf}").argtypes = [vp, vp, C.c_double, C.c_int, C.c_int64, C.c_double,
                                                                  vp, vp, C.POINTER(C.c_double)]
                getattr(L, f"bo_poisson_solve_{suf}").restype = C.c_int64
                getattr(L, f"bo_grav_accel_{suf}").argtypes = [vp, vp, vp, vp]
                getattr(L, f"bo_step_grav_{suf}").argtypes = [vp, vp, vp, C.c_double, vp, C.c_int, vp]
                getattr(L, f"bo_step_grav_{suf}").restype = C.c_int
                getattr(L, f"bo_compute_dt_grav_{suf}").argtypes = [vp, vp, C.c_double, C.c_int, C.c_double]
                getattr(L, f"bo_compute_dt_grav_{suf}").restype = C.c_double
                getattr(L, f"bo_evolve_grav_{suf}").argtypes = [vp, vp, vp, C.c_double, C.c_double, C.c_int64,
                                                                C.c_int, C.c_int, C.c_double, C.c_int, C.c_int64,
                                                                C.c_double, vp, C.POINTER(C.c_double),
                                                                C.POINTER(C.c_int64), C.POINTER(C.c_double)]
                getattr(L, f"bo_evolve_grav_{suf}").restype = C.c_int64
            L._bo_typed = True
        self.L = L
        self.mesh = mesh
        self._keep = [np.ascontiguousarray(mesh.level, dtype=np.int32),
                      np.ascontiguousarray(mesh.dims, dtype=np.int32),
                      np.ascontiguousarray(mesh.origin, dtype=np.int64),
                      np.ascontiguousarray(mesh.bc, dtype=np.int32)]
        h0 = (C.c_double * 3)(*mesh.h0[:3])
        self.h = L.bo_create(mesh.rank, len(mesh.level), mesh.maxlevel, h0, mesh.gamma,
                             *[a.ctypes.data for a in self._keep])
        if not self.h:
            raise ValueError("invalid block mesh")
        self.ncell = int(L.bo_ncell(self.h))

    def __del__(self):
        if getattr(self, "_poisson", None):
            self.L.bo_poisson_destroy(self._poisson)
            self._poisson = None
        if getattr(self, "h", None):
            self.L.bo_destroy(self.h)
            self.h = None

    # ---- self-gravity on the block mesh (blocks_gravity.c) ----
    def _P(self):
        if not getattr(self, "_poisson", None):
            self._poisson = self.L.bo_poisson_create(self.h)
        return self._poisson

    def poisson_csr(self):
        """(rowptr, colidx, coef, glo, ghi, gsum) of the composite operator and
        the gravity stencil ([cell][axis] for the last three)."""
        P = self._P()
        n, nnz = self.ncell, int(self.L.bo_poisson_nnz(P))
        rp, ci = np.empty(n + 1, np.int64), np.empty(nnz, np.int64)
        cf = np.empty(nnz)
        glo, ghi = np.empty((n, 3), np.int64), np.empty((n, 3), np.int64)
        gs = np.empty((n, 3))
        self.L.bo_poisson_export(P, *[a.ctypes.data for a in (rp, ci, cf, glo, ghi, gs)])
        return rp, ci, cf, glo, ghi, gs

    def poisson_apply(self, x):
        x = np.ascontiguousarray(x)
        out = np.empty_like(x)
        getattr(self.L, f"bo_poisson_apply_{_suf(x.dtype)}")(self._P(), x.ctypes.data, out.ctypes.data)
        return out

    def poisson_rhs(self, U, G, project_mean):
        U = np.ascontiguousarray(U)
        b = np.empty(self.ncell, dtype=U.dtype)
        getattr(self.L, f"bo_poisson_rhs_{_suf(U.dtype)}")(self._P(), U.ctypes.data, float(G), int(project_mean),
                                                           b.ctypes.data)
        return b

    def poisson_solve(self, U, G, project_mean, maxiter, tol, phi):
        """CG from the warm start phi (updated in place); returns (iters, relres)."""
        U = np.ascontiguousarray(U)
        assert phi.flags.c_contiguous and phi.dtype == U.dtype
        work = np.empty(4 * self.ncell, dtype=U.dtype)
        rr = C.c_double()
        it = getattr(self.L, f"bo_poisson_solve_{_suf(U.dtype)}")(self._P(), U.ctypes.data, float(G),
                                                                  int(project_mean), int(maxiter), float(tol),
                                                                  phi.ctypes.data, work.ctypes.data, C.byref(rr))
        return int(it), rr.value

    def grav_accel(self, phi):
        phi = np.ascontiguousarray(phi)
        g = np.empty(3 * self.ncell, dtype=phi.dtype)
        getattr(self.L, f"bo_grav_accel_{_suf(phi.dtype)}")(self.h, self._P(), phi.ctypes.data, g.ctypes.data)
        return g.reshape(3, self.ncell)

    def step_grav(self, U, dt, phi, floor64=False):
        assert U.flags.c_contiguous
        phi = np.ascontiguousarray(phi, dtype=U.dtype)
        work = np.empty(10 * self.ncell, dtype=U.dtype)
        getattr(self.L, f"bo_step_grav_{_suf(U.dtype)}")(self.h, self._P(), U.ctypes.data, float(dt), work.ctypes.data,
                                                         int(floor64), phi.ctypes.data)

    def compute_dt_grav(self, U, cfl, G, floor64=False):
        U = np.ascontiguousarray(U)
        return getattr(self.L, f"bo_compute_dt_grav_{_suf(U.dtype)}")(self.h, U.ctypes.data, float(cfl),
                                                                     int(floor64), float(G))

    def evolve_grav(self, U, tfinal, cfl, G, phi, max_steps=0, subcycle=False, floor64=False, t0=0.0,
                    project_mean=True, maxiter=2000, tol=1e-10):
        """evolve! with self-gravity (blocks_gravity_impl.h bo_evolve_grav): a
        Poisson solve before each root step, single rate or subcycled. U and
        phi are updated in place → (t, root steps, CG iterations, worst relres)."""
        assert U.flags.c_contiguous and phi.flags.c_contiguous and phi.dtype == U.dtype
        t = C.c_double(t0)
        it = C.c_int64()
        rr = C.c_double()
        k = getattr(self.L, f"bo_evolve_grav_{_suf(U.dtype)}")(
            self.h, self._P(), U.ctypes.data, float(tfinal), float(cfl), int(max_steps), int(subcycle), int(floor64),
            float(G), int(project_mean), int(maxiter), float(tol), phi.ctypes.data, C.byref(t), C.byref(it),
            C.byref(rr))
        return t.value, int(k), int(it.value), rr.value

    def validate(self) -> int:
        return int(self.L.bo_validate(self.h))

    def accumulate_flux(self, U, floor64=False):
        U = np.ascontiguousarray(U)
        acc = np.zeros_like(U)
        getattr(self.L, f"bo_accumulate_flux_{_suf(U.dtype)}")(self.h, U.ctypes.data, acc.ctypes.data, int(floor64))
        return acc

    def step(self, U, dt, level=-1, floor64=False):
        assert U.flags.c_contiguous
        work = np.empty(10 * self.ncell, dtype=U.dtype)
        getattr(self.L, f"bo_step_{_suf(U.dtype)}")(self.h, U.ctypes.data, float(dt), int(level), work.ctypes.data,
                                                   int(floor64))

    def compute_dt(self, U, cfl, level=-1, floor64=False):
        U = np.ascontiguousarray(U)
        return getattr(self.L, f"bo_compute_dt_{_suf(U.dtype)}")(self.h, U.ctypes.data, cfl, int(level),
                                                                int(floor64))

    def totals(self, U):
        U = np.ascontiguousarray(U)
        out = np.zeros(5)
        getattr(self.L, f"bo_totals_{_suf(U.dtype)}")(self.h, U.ctypes.data, out.ctypes.data)
        return out

    def indicator(self, U):
        """density_gradient_indicator per leaf and its per-block maximum
        (blocks_impl.h bo_indicator) → (block_max, cell_g)."""
        U = np.ascontiguousarray(U)
        cell = np.empty(self.ncell)
        bmax = np.empty(len(self.mesh.level))
        getattr(self.L, f"bo_indicator_{_suf(U.dtype)}")(self.h, U.ctypes.data, cell.ctypes.data, bmax.ctypes.data)
        return bmax, cell

    def evolve(self, U, tfinal, cfl, max_steps=0, subcycle=False, floor64=False, t0=0.0):
        assert U.flags.c_contiguous
        t = C.c_double(t0)
        k = getattr(self.L, f"bo_evolve_{_suf(U.dtype)}")(self.h, U.ctypes.data, tfinal, cfl, int(max_steps),
                                                         int(subcycle), int(floor64), C.byref(t))
        return t.value, int(k)


def blocks_remap(dst, src, src_block, U_src: np.ndarray) -> np.ndarray:
    """Injection remap between two block meshes (prolong!,
    lib/RefMesh/src/RefMesh.jl:284-302): every cell of block b of `dst` takes
    the fields of the cell of block src_block[b] of `src` that contains it.
    `dst`/`src` are BlockMesh-like (rank, level, dims, origin, maxlevel)."""
    def offsets(m):
        return np.concatenate([[0], np.cumsum(np.prod(np.asarray(m.dims, dtype=np.int64), axis=1))])
    so, do = offsets(src), offsets(dst)
    nf = U_src.shape[0]
    out = np.empty((nf, int(do[-1])), dtype=U_src.dtype)
    for b in range(len(dst.level)):
        p = int(src_block[b])
        shift = int(dst.level[b]) - int(src.level[p])
        if shift < 0:
            raise ValueError("destination block coarser than its source")
        sspan = 1 << (src.maxlevel - int(src.level[p]))
        nd = [int(v) for v in dst.dims[b]]
        ns = [int(v) for v in src.dims[p]]
        idx = []
        for d in range(3):
            if d >= dst.rank:
                idx.append(np.zeros(nd[d], dtype=np.int64))
                continue
            rel = int(dst.origin[b][d]) - int(src.origin[p][d])
            if rel < 0 or rel % sspan:
                raise ValueError("destination block not aligned to its source")
            ii = rel // sspan + (np.arange(nd[d]) >> shift)
            if ii[-1] >= ns[d]:
                raise ValueError("destination block not inside its source")
            idx.append(ii)
        K, J, I = np.meshgrid(idx[2], idx[1], idx[0], indexing="ij")
        sc = so[p] + (K * ns[1] + J) * ns[0] + I
        out[:, do[b]:do[b + 1]] = U_src[:, sc.ravel()]
    return out


# ---------------- self-gravity (src/gravity.jl) ----------------

def gper_mask(rank: int, bcs) -> int:
    """Bit mask of the axes along which φ is periodic; bcs: "periodic"/"neumann"
    for every axis, or one name per axis."""
    if isinstance(bcs, str):
        bcs = [bcs] * rank
    return sum(1 << d for d in range(rank) if bcs[d] == "periodic")


def poisson_apply(g: Grid, gper: int, x: np.ndarray) -> np.ndarray:
    """A·x with A = −∇²·V (src/gravity.jl:99-128, 326-336)."""
    x = np.ascontiguousarray(x)
    out = np.empty_like(x)
    getattr(lib(), f"vo_poisson_apply_{_suf(x.dtype)}")(C.byref(g), int(gper), _ptr(x), _ptr(out))
    return out


def poisson_rhs(g: Grid, U: np.ndarray, G: float, project_mean: bool) -> np.ndarray:
    U = np.ascontiguousarray(U)
    b = np.empty(ncell(g), dtype=U.dtype)
    getattr(lib(), f"vo_poisson_rhs_{_suf(U.dtype)}")(C.byref(g), _ptr(U), float(G), int(project_mean), _ptr(b))
    return b


def poisson_solve(g: Grid, gper: int, U: np.ndarray, G: float, project_mean: bool, maxiter: int, tol: float,
                  phi: np.ndarray):
    """CG solve of ∇²φ = 4πGρ into phi (warm start, in place); returns (iters, relres)."""
    U = np.ascontiguousarray(U)
    assert phi.flags.c_contiguous and phi.dtype == U.dtype
    work = np.empty(4 * ncell(g), dtype=U.dtype)
    rr = C.c_double()
    it = getattr(lib(), f"vo_poisson_solve_{_suf(U.dtype)}")(C.byref(g), int(gper), _ptr(U), float(G),
                                                            int(project_mean), int(maxiter), float(tol),
                                                            _ptr(phi), _ptr(work), C.byref(rr))
    return int(it), rr.value


def grav_accel(g: Grid, gper: int, phi: np.ndarray) -> np.ndarray:
    phi = np.ascontiguousarray(phi)
    out = np.empty(3 * ncell(g), dtype=phi.dtype)
    getattr(lib(), f"vo_grav_accel_{_suf(phi.dtype)}")(C.byref(g), int(gper), _ptr(phi), _ptr(out))
    return out.reshape(3, -1)


def step_grav(g: Grid, U: np.ndarray, dt: float, gper: int, phi: np.ndarray, floor64=False) -> None:
    """In-place SSP-RK2 step with the gravity source of the given φ."""
    assert U.flags.c_contiguous
    phi = np.ascontiguousarray(phi, dtype=U.dtype)
    work = np.empty(10 * ncell(g), dtype=U.dtype)
    rc = getattr(lib(), f"vo_step_grav_{_suf(U.dtype)}")(C.byref(g), _ptr(U), float(dt), _ptr(work), int(floor64),
                                                         int(gper), _ptr(phi))
    if rc != 0:
        raise ValueError("invalid grid")


def compute_dt_grav(g: Grid, U: np.ndarray, cfl: float, G: float, floor64=False) -> float:
    U = np.ascontiguousarray(U)
    return float(getattr(lib(), f"vo_compute_dt_grav_{_suf(U.dtype)}")(C.byref(g), _ptr(U), float(cfl),
                                                                       int(floor64), float(G)))


# ---------------- PPMKernels unsplit RK2 MUSCL (muscl_oracle.c) ----------------

def _muscl_typed():
    L = lib()
    if not getattr(L, "_mu_typed", False):
        vp = C.c_void_p
        i64p = C.POINTER(C.c_int64)
        for suf in ("f64", "f32"):
            getattr(L, f"vo_muscl_flux_line_{suf}").argtypes = [vp] * 5 + [C.c_int64] * 3 + [C.c_double] * 3 + [vp] * 6
            getattr(L, f"vo_muscl_step_3d_{suf}").argtypes = [vp, i64p, C.c_int64] + [C.c_double] * 5 + \
                [C.c_int, C.c_double, vp]
            getattr(L, f"vo_muscl_step_3d_bc_{suf}").argtypes = [vp, i64p, C.c_int64] + [C.c_double] * 5 + \
                [C.c_int, C.c_double, vp, vp, vp]
            getattr(L, f"vo_muscl_fill_periodic_{suf}").argtypes = [vp, i64p, C.c_int64]
            getattr(L, f"vo_muscl_max_wavespeed_{suf}").argtypes = [vp, C.c_int64, C.c_double]
            getattr(L, f"vo_muscl_max_wavespeed_{suf}").restype = C.c_double
        L._mu_typed = True
    return L


def muscl_flux_line(rho, eint, vx, vy, vz, ncells, nghost, jdim=1, gamma=1.4, theta=1.5, small_rho=1e-10,
                    ge_flux=False):
    """muscl_flux_line! (lib/PPMKernels/src/muscl.jl:223-239) → (fd, fs1, fs2, fs3, fe[, fge]),
    each of (ncells − 2·nghost + 1)·jdim interfaces."""
    L = _muscl_typed()
    arrs = [np.ascontiguousarray(a) for a in (rho, eint, vx, vy, vz)]
    dt = arrs[0].dtype
    nfi = ncells - 2 * nghost + 1
    out = [np.empty(nfi * jdim, dtype=dt) for _ in range(6 if ge_flux else 5)]
    getattr(L, f"vo_muscl_flux_line_{_suf(dt)}")(*[a.ctypes.data for a in arrs], int(ncells), int(nghost), int(jdim),
                                                  float(gamma), float(theta), float(small_rho),
                                                  *[o.ctypes.data for o in out], out[5].ctypes.data if ge_flux else None)
    return tuple(out)


def muscl_step_3d(U, dims, ng, dt, gamma=1.4, theta=1.5, dx=1.0, small_rho=1e-10, record=False, eta1=1e-3):
    """muscl_step_3d! (lib/PPMKernels/src/muscl_grid.jl:314-370) on U in place: (5, N), or
    (6, N) with the gas energy ρe as the sixth field (dual-energy formalism, eta1);
    with record=True returns the recorded fluxes (3, nf, N) (½F per stage)."""
    L = _muscl_typed()
    assert U.flags.c_contiguous and U.shape[0] in (5, 6)
    nf = U.shape[0]
    d = (C.c_int64 * 3)(*dims)
    frec = np.empty((3, nf, U.shape[1]), dtype=U.dtype) if record else None
    getattr(L, f"vo_muscl_step_3d_{_suf(U.dtype)}")(U.ctypes.data, d, int(ng), float(dt), float(gamma), float(theta),
                                                     float(dx), float(small_rho), nf, float(eta1),
                                                     frec.ctypes.data if record else None)
    return frec


_MU_BC = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_int)


def muscl_step_3d_bc(U, dims, ng, dt, bc, gamma=1.4, theta=1.5, dx=1.0, small_rho=1e-10, record=False, eta1=1e-3):
    """muscl_step_3d! with bc!: bc(V, stage) refills the ghost zones of V, a
    (nf, N) view of the state, before each stage (muscl_impl.h vo_muscl_step_3d_bc)."""
    L = _muscl_typed()
    assert U.flags.c_contiguous and U.shape[0] in (5, 6)
    nf, N = U.shape
    d = (C.c_int64 * 3)(*dims)
    frec = np.empty((3, nf, N), dtype=U.dtype) if record else None
    ct = C.c_double if U.dtype == np.float64 else C.c_float

    def hook(_user, ptr, stage):
        V = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(nf, N))
        bc(V, int(stage))

    cb = _MU_BC(hook)
    getattr(L, f"vo_muscl_step_3d_bc_{_suf(U.dtype)}")(U.ctypes.data, d, int(ng), float(dt), float(gamma),
                                                        float(theta), float(dx), float(small_rho), nf, float(eta1),
                                                        frec.ctypes.data if record else None,
                                                        C.cast(cb, C.c_void_p), None)
    return frec


def muscl_fill_periodic(f, dims, ng):
    """fill_periodic! of one field (lib/PPMKernels/src/muscl_grid.jl:644-670), in place."""
    L = _muscl_typed()
    assert f.flags.c_contiguous
    getattr(L, f"vo_muscl_fill_periodic_{_suf(f.dtype)}")(f.ctypes.data, (C.c_int64 * 3)(*dims), int(ng))


def muscl_max_wavespeed(U, gamma):
    """max_wavespeed (lib/PPMKernels/src/muscl_grid.jl:672-695)."""
    L = _muscl_typed()
    U = np.ascontiguousarray(U)
    return getattr(L, f"vo_muscl_max_wavespeed_{_suf(U.dtype)}")(U.ctypes.data, U.shape[1], float(gamma))
