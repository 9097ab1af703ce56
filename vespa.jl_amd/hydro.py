"""Host-side mirror of Vespa.jl's flux-backend interface over the HIP library.

The reference's device seam is ``compute_face_fluxes!(backend, model, F, WL, WR,
axisv)`` (src/ka_flux.jl:40-53; KA method ext/VespaKAFluxExt.jl:37-63) behind
``sim.flux`` (src/driver.jl:44, dispatch :210-211), and the whole-stage unit is
``step!(sim, dt)`` (src/driver.jl:333-359). This module keeps those names,
argument meanings and error behaviour, with Python's ValueError standing in for
Julia's ArgumentError:

* ``IdealHydro(γ)``                 src/equationset.jl:31-60
* ``HIPFlux()`` + ``compute_face_fluxes``  the AbstractFluxBackend drop-in
* ``Problem``, ``sod_problem``      src/problem.jl:140-191, problems/sod_shock_tube.jl
* ``UniformMesh(dims, domain)``     lib/RefMesh/src/RefMesh.jl:89-142
* ``Simulation``, ``step``, ``compute_dt``, ``evolve``  src/driver.jl:62-79,169-199,333-359,483-508
* ``conserved_totals``, ``dump_fields``  src/diagnostics.jl:16-31,248-258

The conserved state lives on the GPU inside a ``vh_grid`` context; the host only
uploads the initial condition and downloads on request.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Callable, Sequence

import numpy as np

from . import _lib
from ._lib import check, load_library

# ---------------------------------------------------------------- models / BCs


@dataclass(frozen=True)
class IdealHydro:
    """Adiabatic ideal gas, conserved (ρ, ρvx, ρvy, ρvz, E), HLLC flux."""
    gamma: float

    nvars = 5
    conserved_names = ("density", "momentum_x", "momentum_y", "momentum_z", "total_energy")
    density_index = 1
    momentum_indices = (2, 3, 4)
    energy_index = 5

    def adiabatic_index(self) -> float:
        return self.gamma


class Outflow:
    code = _lib.BC_OUTFLOW


class Periodic:
    code = _lib.BC_PERIODIC


class Reflecting:
    code = _lib.BC_REFLECTING


class ParentGhost:
    """ParentGhost(ghost) (lib/MeshInterface/src/MeshInterface.jl:86-99): the
    boundary ghost of a face on this side is ghost(axis, side, cell), a
    primitive state (ρ, vx, vy, vz, p) — axis 1-based, side "lo" or "hi", cell
    the 1-based index tuple of the boundary cell — used as the outer neighbour
    of the cell's slope and as the outer state of its boundary face
    (src/driver.jl:127-134). The Simulation evaluates it for every face of the
    side before each step and hands the states to the device."""
    code = _lib.BC_PARENT

    def __init__(self, ghost):
        self.ghost = ghost


def _bc_code(bc) -> int:
    if isinstance(bc, str):
        return {"periodic": _lib.BC_PERIODIC, "outflow": _lib.BC_OUTFLOW,
                "reflecting": _lib.BC_REFLECTING}[bc.lower()]
    if isinstance(bc, type) and hasattr(bc, "code"):
        return bc.code
    if hasattr(bc, "code"):
        return bc.code
    raise ValueError(f"unknown boundary condition {bc!r}")


def _parent_sides(bcs, rank: int) -> dict:
    """{(axis 0-based, side 0/1): ParentGhost} of the sides whose BC is one."""
    if isinstance(bcs, (list, tuple)) and len(bcs) == rank and all(
            isinstance(p, (list, tuple)) and len(p) == 2 for p in bcs):
        pairs = list(bcs)
    else:
        pairs = [(bcs, bcs)] * rank
    return {(d, s): pairs[d][s] for d in range(rank) for s in range(2) if isinstance(pairs[d][s], ParentGhost)}


def _bcs_array(bcs, rank: int):
    if isinstance(bcs, (list, tuple)) and len(bcs) == rank and all(
            isinstance(p, (list, tuple)) and len(p) == 2 for p in bcs):
        pairs = [(_bc_code(a), _bc_code(b)) for a, b in bcs]
    else:
        c = _bc_code(bcs)
        pairs = [(c, c)] * rank
    arr = (C.c_int32 * 6)()
    for d in range(3):
        lo, hi = pairs[d] if d < rank else (_lib.BC_OUTFLOW, _lib.BC_OUTFLOW)
        arr[2 * d], arr[2 * d + 1] = lo, hi
    return arr


# ---------------------------------------------------------------- tier 1


class HIPFlux:
    """AbstractFluxBackend running the per-face HLLC on the GPU (the KAFlux drop-in)."""

    def compute_face_fluxes(self, model, F, WL, WR, axisv):
        return compute_face_fluxes(self, model, F, WL, WR, axisv)


def compute_face_fluxes(backend, model, F: np.ndarray, WL: np.ndarray, WR: np.ndarray,
                        axisv: np.ndarray) -> np.ndarray:
    """Fill F[f] with the HLLC flux of (WL[f], WR[f]) across axis axisv[f] ∈ {1,2,3}.

    F, WL, WR are (nf, 5) arrays of one float dtype (the packed NTuple{5,T}
    columns); returns F. Mirrors src/ka_flux.jl:40-53 / ext/VespaKAFluxExt.jl:37-63.
    """
    if not isinstance(backend, HIPFlux):
        raise TypeError(f"no compute_face_fluxes method for {type(backend).__name__}")
    WL = np.asarray(WL)
    dtype = WL.dtype
    if dtype not in (np.float64, np.float32):
        raise TypeError(f"face states must be float64 or float32, got {dtype}")
    nf = WL.shape[0] if WL.ndim else 0
    if WL.shape != (nf, 5) or np.shape(WR) != (nf, 5) or np.shape(F) != (nf, 5):
        raise ValueError("WL, WR and F must all be (nf, 5)")
    if np.shape(axisv) != (nf,):
        raise ValueError("axisv must have length nf")
    if nf == 0:
        return F
    wl = np.ascontiguousarray(WL)
    wr = np.ascontiguousarray(WR, dtype=dtype)
    ax = np.ascontiguousarray(axisv, dtype=np.int32)
    if ax.min() < 1 or ax.max() > 3:
        raise ValueError("axisv entries must be 1, 2 or 3")
    out = F if (isinstance(F, np.ndarray) and F.flags.c_contiguous and F.dtype == dtype) else \
        np.empty((nf, 5), dtype=dtype)
    L = load_library()
    tp = C.c_double if dtype == np.float64 else C.c_float
    fn = L.vh_face_flux_f64 if dtype == np.float64 else L.vh_face_flux_f32
    check(fn(wl.ctypes.data_as(C.POINTER(tp)), wr.ctypes.data_as(C.POINTER(tp)),
             ax.ctypes.data_as(C.POINTER(C.c_int32)), nf, float(model.adiabatic_index()),
             out.ctypes.data_as(C.POINTER(tp))))
    if out is not F:
        F[...] = out
    return F


# ---------------------------------------------------------------- problem / mesh


@dataclass
class UniformMesh:
    """Single-level uniform mesh: dims per axis, domain = ((lo, hi), ...)."""
    dims: tuple
    domain: tuple

    def __post_init__(self):
        self.dims = tuple(int(d) for d in self.dims)
        self.domain = tuple((float(a), float(b)) for a, b in self.domain)
        if not 1 <= len(self.dims) <= 3 or len(self.domain) != len(self.dims):
            raise ValueError("UniformMesh needs 1..3 dims with one (lo, hi) per axis")

    @property
    def rank(self) -> int:
        return len(self.dims)

    def cell_width(self):
        return tuple((hi - lo) / n for (lo, hi), n in zip(self.domain, self.dims))

    def cell_volume(self) -> float:
        w = self.cell_width()
        v = w[0]
        for x in w[1:]:
            v = v * x
        return v

    def cell_centers(self):
        """Per-axis centre coordinates lo + (I − ½)·w, I = 1..n (RefMesh.jl:129-132)."""
        w = self.cell_width()
        return [lo + (np.arange(1, n + 1, dtype=np.float64) - 0.5) * wd
                for (lo, _), n, wd in zip(self.domain, self.dims, w)]


@dataclass
class Problem:
    """A hydro problem (src/problem.jl:140-170). ``init(x, y, z)`` receives numpy
    coordinate arrays (broadcast over the mesh; absent axes are 0.0) and returns
    the primitive tuple (ρ, vx, vy, vz, p) as arrays or scalars."""
    name: str
    dims: tuple
    domain: tuple
    gamma: float
    bcs: object
    init: Callable
    tfinal: float
    cfl: float = 0.4

    def __post_init__(self):
        if not self.gamma > 1:
            raise ValueError(f"γ must exceed 1 (got {self.gamma})")
        if not self.tfinal > 0:
            raise ValueError(f"tfinal must be positive (got {self.tfinal})")
        if not 0 < self.cfl <= 1:
            raise ValueError(f"cfl must be in (0, 1] (got {self.cfl})")
        if not all(d > 0 for d in self.dims):
            raise ValueError("all dims must be positive")
        for d, (lo, hi) in enumerate(self.domain):
            if not hi > lo:
                raise ValueError(f"domain[{d + 1}] must have hi > lo (got {(lo, hi)})")


def sod_problem(n: int = 128) -> Problem:
    """Sod shock tube (problems/sod_shock_tube.jl:223-234)."""
    def init(x, y, z):
        left = x < 0.5
        return (np.where(left, 1.0, 0.125), 0.0, 0.0, 0.0, np.where(left, 1.0, 0.1))
    return Problem("SodShockTube", (int(n),), ((0.0, 1.0),), 1.4, "outflow", init, 0.2, 0.4)


def prim2cons_array(W: np.ndarray, gamma: float, dtype=np.float64) -> np.ndarray:
    """prim2cons (src/eos.jl:28-32) over a (5, ...) primitive array, in `dtype`."""
    W = np.asarray(W, dtype=dtype)
    T = np.dtype(dtype).type
    g = T(gamma)
    r, vx, vy, vz, p = W
    E = p / (g - T(1)) + (T(0.5) * r) * ((vx * vx + vy * vy) + vz * vz)
    return np.stack([r, r * vx, r * vy, r * vz, E]).astype(dtype, copy=False)


def cons2prim_array(U: np.ndarray, gamma: float) -> np.ndarray:
    """cons2prim (src/eos.jl:8-25) over a (5, ...) conserved array in its own dtype."""
    U = np.asarray(U)
    T = U.dtype.type
    g = T(gamma)
    r = U[0]
    inv = T(1) / r
    vx, vy, vz = U[1] * inv, U[2] * inv, U[3] * inv
    kin = (T(0.5) * r) * ((vx * vx + vy * vy) + vz * vz)
    p = (g - T(1)) * (U[4] - kin)
    p = np.maximum(p, T(1e-12) * kin)
    return np.stack([r, vx, vy, vz, p])


# ---------------------------------------------------------------- simulation


class Simulation:
    """Device-resident simulation on a UniformMesh (src/driver.jl:27-79).

    ``eltype`` is the field precision Tf (np.float64 or np.float32); geometry,
    the clock and the reductions stay Float64 (the decoupled split,
    src/driver.jl:48-52)."""

    def __init__(self, mesh: UniformMesh, prob: Problem, eltype=np.float64, flux=None):
        self.mesh = mesh
        self.problem = prob
        self.model = IdealHydro(prob.gamma)
        self.eltype = np.dtype(eltype)
        if self.eltype not in (np.float64, np.float32):
            raise ValueError(f"eltype must be float64 or float32, got {self.eltype}")
        # sim.flux (src/driver.jl:44): None runs the fused device stage kernels; a
        # flux backend (HIPFlux()) runs accumulate_flux! in the reference's batched
        # form through it (src/ka_flux.jl:128-162; batched.step_batched).
        if flux is not None and not isinstance(flux, HIPFlux):
            raise TypeError(f"no compute_face_fluxes method for {type(flux).__name__}")
        self.flux = flux
        self.t = 0.0
        self.step_count = 0
        self.grav = None   # GravityField when self-gravity is on (gravity.enable_gravity)
        self._lib = load_library()
        rank = mesh.rank
        dims = (C.c_int64 * 3)(*[mesh.dims[d] if d < rank else 1 for d in range(3)])
        lo = (C.c_double * 3)(*[mesh.domain[d][0] if d < rank else 0.0 for d in range(3)])
        hi = (C.c_double * 3)(*[mesh.domain[d][1] if d < rank else 1.0 for d in range(3)])
        bc = _bcs_array(prob.bcs, rank)
        self._bc_pairs = [(int(bc[2 * d]), int(bc[2 * d + 1])) for d in range(rank)]
        self._parent = _parent_sides(prob.bcs, rank)   # {(axis, side): ParentGhost}
        h = C.c_void_p()
        check(self._lib.vh_grid_create(C.byref(h), rank, dims, lo, hi, bc, float(prob.gamma),
                                       _lib.VH_F64 if self.eltype == np.float64 else _lib.VH_F32))
        self._h = h
        self.ncell = int(np.prod(mesh.dims))
        self.set_initial_conditions()

    # -- lifecycle
    def close(self):
        if getattr(self, "_h", None):
            self._lib.vh_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        """numpy shape of the state: (5, nz, ny, nx) trimmed to the rank."""
        return (5,) + tuple(reversed(self.mesh.dims))

    def set_initial_conditions(self):
        """src/driver.jl:101-112: U = prim2cons(init(centre)) in Tf."""
        cs = self.mesh.cell_centers()
        rank = self.mesh.rank
        grids = np.meshgrid(*reversed(cs), indexing="ij") if rank > 1 else [cs[0]]
        coords = list(reversed(grids)) + [np.zeros(()) for _ in range(3 - rank)]
        W = self.problem.init(*coords)
        shape = tuple(reversed(self.mesh.dims))
        Wa = np.stack([np.broadcast_to(np.asarray(w, dtype=np.float64), shape) for w in W])
        Wt = Wa.astype(self.eltype)        # map(Tf, init(coords...))
        self.upload(prim2cons_array(Wt, self.problem.gamma, self.eltype))

    def upload(self, U: np.ndarray):
        U = np.ascontiguousarray(U, dtype=self.eltype)
        if U.size != 5 * self.ncell:
            raise ValueError(f"state must hold 5*{self.ncell} values")
        check(self._lib.vh_grid_upload(self._h, U.ctypes.data_as(C.c_void_p)))

    def state(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.eltype)
        check(self._lib.vh_grid_download(self._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def device_state_ptr(self) -> int:
        p = C.c_void_p()
        check(self._lib.vh_grid_state_device_ptr(self._h, C.byref(p)))
        return int(p.value)

    def set_mode(self, mode: str):
        """'exact' (bit-identical to the reference arithmetic) or 'fast'."""
        code = {"exact": _lib.VH_EXACT, "fast": _lib.VH_FAST}[mode]
        check(self._lib.vh_grid_set_mode(self._h, code))

    @property
    def mode(self) -> str:
        return {_lib.VH_EXACT: "exact", _lib.VH_FAST: "fast"}[self._lib.vh_grid_get_mode(self._h)]

    def set_launch(self, tile_variant: int = 0, zchunk: int = 0):
        check(self._lib.vh_grid_set_launch(self._h, int(tile_variant), int(zchunk)))

    def launch(self):
        tx, ty, zc, wg = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int64()
        check(self._lib.vh_grid_get_launch(self._h, C.byref(tx), C.byref(ty), C.byref(zc), C.byref(wg)))
        return {"tile": (tx.value, ty.value), "zchunk": zc.value, "workgroups": wg.value}

    def tile_count(self) -> int:
        return int(self._lib.vh_grid_tile_count(self._h))

    def sync(self):
        check(self._lib.vh_grid_sync(self._h))

    def time_steps(self, nsteps: int) -> dict:
        """Run nsteps device-dt steps with HIP events around each stage kernel on
        the context's stream (one synchronisation at the end). Returns mean
        milliseconds per stage-1 and stage-2 launch and per whole step."""
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        check(self._lib.vh_grid_time_steps(self._h, float(self.problem.cfl), int(nsteps),
                                           C.byref(a), C.byref(b), C.byref(c)))
        self.step_count += int(nsteps)
        return {"stage1_ms": a.value, "stage2_ms": b.value, "step_ms": c.value}


def compute_dt(sim: Simulation) -> float:
    """CFL time step (src/driver.jl:169-199)."""
    dt = C.c_double()
    check(sim._lib.vh_grid_compute_dt(sim._h, float(sim.problem.cfl), C.byref(dt)))
    return dt.value


def parent_ghost_states(sim: Simulation) -> np.ndarray:
    """The ParentGhost closures of sim evaluated for every boundary face of their
    sides: (bflux_count, 5) primitives at the state's precision in the
    boundary-face slot order (other planes zero)."""
    dims = tuple(sim.mesh.dims)
    rank = len(dims)
    n = int(np.prod(dims))
    out = np.zeros((bflux_count(sim), 5), dtype=sim.eltype)
    for (d, side), pg in sorted(sim._parent.items()):
        base = sum(2 * (n // dims[e]) for e in range(d)) + (n // dims[d] if side else 0)
        other = [e for e in range(rank) if e != d]
        shape = [dims[e] for e in other]
        for q in range(n // dims[d]):
            cell = [0] * rank
            cell[d] = 1 if side == 0 else dims[d]
            r = q
            for e, ne in zip(other, shape):
                cell[e] = r % ne + 1
                r //= ne
            out[base + q] = np.asarray(pg.ghost(d + 1, "hi" if side else "lo", tuple(cell)), dtype=np.float64)
    return out


def set_parent_ghost(sim: Simulation, states: np.ndarray) -> None:
    """Hand ParentGhost states (bflux_count × 5 primitives, boundary-face slot
    order) to the device (vh_grid_set_parent_ghost)."""
    a = np.ascontiguousarray(states, dtype=sim.eltype)
    if a.shape != (bflux_count(sim), 5):
        raise ValueError(f"states must be ({bflux_count(sim)}, 5)")
    check(sim._lib.vh_grid_set_parent_ghost(sim._h, a.ctypes.data_as(C.c_void_p)))


def _refresh_parent(sim: Simulation) -> None:
    if sim._parent:
        set_parent_ghost(sim, parent_ghost_states(sim))


def step(sim: Simulation, dt: float, bflux: "BoundaryFluxRegister | None" = None) -> Simulation:
    """One SSP-RK2 step (src/driver.jl:333-359); advances sim.t and the step
    count. With a BoundaryFluxRegister, ½dt·F·area of every domain-boundary face
    is added to it per RK stage (src/reflux.jl:105-132). ParentGhost sides take
    their closures' states, evaluated before the step."""
    if sim._parent and sim.flux is not None:
        raise ValueError("ParentGhost boundaries run in the fused device path (sim.flux = None)")
    _refresh_parent(sim)
    if sim.flux is not None:
        if bflux is not None:
            raise ValueError("the boundary-flux register is recorded by the fused device path (sim.flux = None)")
        if sim.grav is not None:
            raise ValueError("self-gravity runs in the fused device path (sim.flux = None)")
        from .batched import step_batched
        step_batched(sim, dt)
    elif bflux is None:
        check(sim._lib.vh_grid_step(sim._h, float(dt)))
    else:
        if bflux.flux.dtype != sim.eltype or bflux.flux.shape != (bflux_count(sim), 5):
            raise ValueError("register does not match this simulation")
        buf = _lib.DeviceBuffer(bflux.flux.nbytes).upload(bflux.flux)
        check(sim._lib.vh_grid_step_bflux(sim._h, float(dt), buf.ptr))
        buf.download(bflux.flux)
        buf.free()
    sim.t += dt
    sim.step_count += 1
    return sim


def bflux_count(sim: Simulation) -> int:
    return int(sim._lib.vh_grid_bflux_count(sim._h))


class BoundaryFluxRegister:
    """Time-integrated flux ∫F·area dt through each domain-boundary face
    (src/reflux.jl:105-122), dense: one row of 5 values per boundary face slot,
    planes ordered (axis 1 lo, axis 1 hi, axis 2 lo, ...), each column-major over
    the other axes (dim 1 fastest)."""

    def __init__(self, sim: Simulation):
        self.dims = tuple(sim.mesh.dims)
        self.flux = np.zeros((bflux_count(sim), 5), dtype=sim.eltype)

    def plane(self, axis: int, side: str) -> np.ndarray:
        """The (axis, side) plane (axis 1-based, side :lo/:hi as "lo"/"hi") as an
        array shaped like the face plane (other axes, dim 1 fastest) × 5."""
        rank = len(self.dims)
        n = int(np.prod(self.dims))
        base = sum(2 * (n // self.dims[e]) for e in range(axis - 1))
        size = n // self.dims[axis - 1]
        if side == "hi":
            base += size
        other = [self.dims[e] for e in range(rank) if e != axis - 1]
        return self.flux[base:base + size].reshape(tuple(reversed(other)) + (5,) if other else (5,))

    def lookup(self, axis: int, side: str, cell) -> np.ndarray:
        """Recorded flux of the boundary face of `cell` (1-based CartesianIndex
        tuple) on (axis, side)."""
        P = self.plane(axis, side)
        idx = tuple(int(cell[e]) - 1 for e in range(len(self.dims)) if e != axis - 1)
        return P[tuple(reversed(idx))] if idx else P


def bflux_plane(rank: int, dim: int, start, stop, g0, flux_off: int, comp: int, Vcell: float, lookup) -> np.ndarray:
    """Rasterize one (dim, side) face plane in Enzo's flat layout
    (src/reflux.jl:169-191): column-major over the orthogonal dims (dim 0
    fastest), the flux dim collapsed; each entry is the recorded flux of
    component `comp` (1-based, 0 → zeros) divided by Vcell. `lookup(I)` returns
    the recorded 5-vector for the active cell I (1-based tuple) or None."""
    Dim = [stop[d] - start[d] + 1 for d in range(rank)]
    n = int(np.prod(Dim))
    out = np.empty(n)
    for lin in range(n):
        rem = lin
        active = []
        for d in range(rank):
            od = rem % Dim[d]
            rem //= Dim[d]
            g = start[d] + od
            active.append(flux_off if d == dim else g - g0[d] + 1)
        F = lookup(tuple(active))
        out[lin] = float(F[comp - 1]) / float(Vcell) if (F is not None and comp > 0) else 0.0
    return out


def run_steps(sim: Simulation, nsteps: int) -> Simulation:
    """nsteps steps with the CFL dt reduced on the device before each one (no
    clip, no host synchronisation inside the loop; with ParentGhost sides the
    closures are evaluated before each step)."""
    if sim.flux is not None or sim._parent:
        for _ in range(int(nsteps)):
            step(sim, compute_dt(sim))
        return sim
    t = C.c_double(sim.t)
    k = C.c_int64()
    check(sim._lib.vh_grid_run(sim._h, float(sim.problem.cfl), 0.0, int(nsteps), C.byref(t), C.byref(k)))
    sim.t = t.value
    sim.step_count += k.value
    return sim


def evolve(sim: Simulation, max_steps: int = 0, callback=None, callback_every: int = 1) -> Simulation:
    """Integrate to problem.tfinal (src/driver.jl:483-508): dt = min(compute_dt,
    tfinal − t), final step clipped to land on tfinal."""
    if callback is not None:
        callback(sim, "init")
    if callback is None and sim.flux is None and not sim._parent:
        t = C.c_double(sim.t)
        k = C.c_int64()
        check(sim._lib.vh_grid_run(sim._h, float(sim.problem.cfl), float(sim.problem.tfinal),
                                   int(max_steps), C.byref(t), C.byref(k)))
        sim.t = t.value
        sim.step_count += k.value
    else:
        tf = sim.problem.tfinal
        n = 0
        while sim.t < tf * (1 - 1e-12) and (max_steps <= 0 or n < max_steps):
            if sim.grav is not None:
                sim.grav.solve()   # φ(ρⁿ), held over the step (src/driver.jl:498)
            dt = min(compute_dt(sim), tf - sim.t)
            step(sim, dt)
            n += 1
            if callback is not None and callback_every > 0 and sim.step_count % callback_every == 0:
                callback(sim, "step")
        sim.sync()
    if callback is not None:
        callback(sim, "final")
    return sim


def conserved_totals(sim: Simulation) -> dict:
    """Σ U·V per conserved field in Float64 (src/diagnostics.jl:16-31)."""
    out = (C.c_double * 5)()
    check(sim._lib.vh_grid_totals(sim._h, out))
    return {"mass": out[0], "momentum_x": out[1], "momentum_y": out[2], "momentum_z": out[3],
            "energy": out[4]}


def dump_fields(sim: Simulation) -> dict:
    """1D primitive profiles sorted by x (src/diagnostics.jl:248-258)."""
    if sim.mesh.rank != 1:
        raise ValueError("dump_fields currently supports 1D; use state() for ND")
    U = sim.state()
    W = cons2prim_array(U, sim.problem.gamma)
    return {"x": sim.mesh.cell_centers()[0], "density": W[0], "velocity_x": W[1], "pressure": W[4]}
