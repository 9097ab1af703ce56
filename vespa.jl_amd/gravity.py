"""Self-gravity on a uniform-grid Simulation (host mirror of src/gravity.jl).

``enable_gravity(sim, G=1.0, bcs="periodic", ...)`` is the reference's
``enable_gravity!`` (src/gravity.jl:63-80): it attaches a :class:`GravityField`
to the simulation. From then on every RK stage adds the gravity source to the
flux sum on the GPU (apply_gravity_source!, src/gravity.jl:412-447),
``compute_dt`` folds in the free-fall limiter (src/gravity.jl:449-452) and
``evolve`` / ``run_steps`` solve the Poisson equation before each step
(src/driver.jl:498). φ lives on the device; the CG solve (A = −∇²·V, Float64 dot
products) runs there too.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DeviceBuffer, check


def _mask(rank: int, bcs) -> int:
    if isinstance(bcs, str):
        bcs = [bcs] * rank
    if len(bcs) != rank:
        raise ValueError("one φ boundary condition per axis")
    m = 0
    for d, b in enumerate(bcs):
        name = getattr(b, "name", b)
        if name not in ("periodic", "neumann", "outflow", "reflecting"):
            raise ValueError(f"unknown φ boundary condition {b!r}")
        if name == "periodic":
            m |= 1 << d
    return m


class GravityField:
    """φ and the solver settings of one simulation (src/gravity.jl:30-45).
    Non-periodic φ sides are homogeneous Neumann (src/gravity.jl:124-128)."""

    def __init__(self, sim, G: float = 1.0, bcs="periodic", maxiter: int = 500, tol: float | None = None,
                 project_mean: bool | None = None):
        self.sim = sim
        rank = sim.mesh.rank
        self.gper = _mask(rank, bcs)
        self.G = float(G)
        self.maxiter = int(maxiter)
        self.tol = float(tol) if tol is not None else (1e-5 if sim.eltype == np.float32 else 1e-8)
        self.project_mean = bool(self.gper == (1 << rank) - 1) if project_mean is None else bool(project_mean)
        check(sim._lib.vh_grid_enable_gravity(sim._h, self.G, self.gper, self.maxiter, self.tol,
                                              int(self.project_mean)))
        self.last = (0, 0.0)

    def solve(self):
        """solve_poisson! (src/gravity.jl:248-322) for the current density:
        returns (iterations, relative residual)."""
        it, rr = C.c_int64(), C.c_double()
        check(self.sim._lib.vh_grid_solve_poisson(self.sim._h, C.byref(it), C.byref(rr)))
        self.last = (it.value, rr.value)
        return self.last

    def phi(self) -> np.ndarray:
        out = np.empty(tuple(reversed(self.sim.mesh.dims)), dtype=self.sim.eltype)
        check(self.sim._lib.vh_grid_phi_download(self.sim._h, out.ctypes.data_as(C.c_void_p)))
        return out

    def set_phi(self, phi: np.ndarray):
        a = np.ascontiguousarray(phi, dtype=self.sim.eltype)
        if a.size != self.sim.ncell:
            raise ValueError("φ must hold one value per cell")
        check(self.sim._lib.vh_grid_phi_upload(self.sim._h, a.ctypes.data_as(C.c_void_p)))

    def apply_laplacian(self, x: np.ndarray) -> np.ndarray:
        """A·x with A = −∇²·V (apply_laplacian!, src/gravity.jl:99-128)."""
        a = np.ascontiguousarray(x, dtype=self.sim.eltype)
        if a.size != self.sim.ncell:
            raise ValueError("one value per cell")
        dx, do = DeviceBuffer(a.nbytes).upload(a), DeviceBuffer(a.nbytes)
        check(self.sim._lib.vh_grid_apply_laplacian(self.sim._h, dx.ptr, do.ptr))
        out = np.empty_like(a)
        do.download(out)
        dx.free()
        do.free()
        return out

    def accel(self) -> np.ndarray:
        """g = −∇φ per cell (src/gravity.jl:367-410), shape (3,) + mesh shape."""
        n = self.sim.ncell
        buf = DeviceBuffer(3 * n * self.sim.eltype.itemsize)
        check(self.sim._lib.vh_grid_gravity_accel(self.sim._h, buf.ptr))
        out = np.empty((3,) + tuple(reversed(self.sim.mesh.dims)), dtype=self.sim.eltype)
        buf.download(out)
        buf.free()
        return out

    def close(self):
        if self.sim._h:
            check(self.sim._lib.vh_grid_disable_gravity(self.sim._h))
        if getattr(self.sim, "grav", None) is self:
            self.sim.grav = None


def enable_gravity(sim, G: float = 1.0, bcs="periodic", maxiter: int = 500, tol: float | None = None,
                   project_mean: bool | None = None) -> GravityField:
    """enable_gravity!(sim; G, bcs, maxiter, tol, project_mean) (src/gravity.jl:63-80).
    tol defaults to 1e-8 (Float64) / 1e-5 (Float32); the mean is projected by
    default when φ is periodic on every axis."""
    grav = GravityField(sim, G, bcs, maxiter, tol, project_mean)
    sim.grav = grav
    return grav


def solve_poisson(sim):
    """solve_poisson!(sim, sim.grav) → (iterations, relative residual)."""
    if getattr(sim, "grav", None) is None:
        raise ValueError("self-gravity is not enabled on this simulation")
    return sim.grav.solve()
