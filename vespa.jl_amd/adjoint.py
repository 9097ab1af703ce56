"""Reverse mode of the SSP-RK2 step on the GPU (Float64 uniform grids).

The reference has no derivative of its step; this module is the host side of
what a ChainRulesCore ``rrule`` over ``step!(sim, dt)`` (src/driver.jl:333-359)
would call:

* ``step_vjp(sim, U, dt, wbar)``  the pullback of one step with dt held fixed;
* ``rrule_step(sim, dt)``         (new state, pullback) for the simulation's state;
* ``Rollout(sim, nsteps, every)`` a checkpointed rollout (states kept every
  ``every`` steps, segments recomputed in the backward pass) with event timings
  and device memory.

Non-smooth points take the active branch (minmod's selected argument or zero,
the active Davis min/max argument, the upwind HLLC branch, the pressure floor
when it is the larger argument).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import DeviceBuffer, check, load_library


def _dev(arr: np.ndarray) -> DeviceBuffer:
    a = np.ascontiguousarray(arr, dtype=np.float64)
    return DeviceBuffer(a.nbytes).upload(a)


def _require_f64(sim):
    if sim.eltype != np.float64:
        raise ValueError("the reverse mode is Float64 only")


def set_form(sim, fused: bool) -> None:
    """Select the kernel form of the reverse mode for this simulation: the
    single tiled kernel (2D, periodic, sizes multiple of 32×16) or the
    three-pass form."""
    _require_f64(sim)
    check(load_library().vh_grid_set_adjoint_form(sim._h, int(bool(fused))))


def step_vjp(sim, U: np.ndarray, dt: float, wbar: np.ndarray) -> np.ndarray:
    """(∂ step(U, dt)/∂U)ᵀ wbar on the GPU for host arrays shaped like sim.state()."""
    _require_f64(sim)
    L = load_library()
    dU, dW = _dev(U), _dev(wbar)
    dO = DeviceBuffer(dU.nbytes)
    check(L.vh_grid_step_vjp(sim._h, dU.ptr, float(dt), dW.ptr, dO.ptr))
    out = np.empty(np.shape(U), dtype=np.float64)
    dO.download(out)
    for b in (dU, dW, dO):
        b.free()
    return out


def rrule_step(sim, dt: float):
    """Advance sim by one step and return (new state, pullback) where
    pullback(wbar) = (∂ new/∂ old)ᵀ wbar."""
    _require_f64(sim)
    U0 = sim.state()
    from .hydro import step
    step(sim, dt)
    U1 = sim.state()

    def pullback(wbar):
        return step_vjp(sim, U0, dt, wbar)
    return U1, pullback


class Rollout:
    """nsteps steps from the simulation's current state with the device CFL dt,
    checkpoints every `every` steps; backward() returns the gradient w.r.t. the
    initial conserved state of a loss whose gradient w.r.t. the final state is
    given."""

    def __init__(self, sim, nsteps: int, every: int = 20):
        _require_f64(sim)
        self.sim = sim
        self.nsteps = int(nsteps)
        self.every = int(every)
        self._lib = load_library()
        h = C.c_void_p()
        check(self._lib.vh_adjoint_create(C.byref(h), sim._h, self.nsteps, self.every))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.vh_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self) -> np.ndarray:
        """Run the rollout; returns the final state (the simulation holds it too)."""
        check(self._lib.vh_adjoint_forward(self._h, float(self.sim.problem.cfl)))
        self.sim.step_count += self.nsteps
        return self.sim.state()

    def backward(self, wbar: np.ndarray) -> np.ndarray:
        dW = _dev(wbar)
        dO = DeviceBuffer(dW.nbytes)
        check(self._lib.vh_adjoint_backward(self._h, dW.ptr, dO.ptr))
        out = np.empty(np.shape(wbar), dtype=np.float64)
        dO.download(out)
        dW.free()
        dO.free()
        return out

    def info(self) -> dict:
        f, b = C.c_double(), C.c_double()
        nbytes = C.c_int64()
        dts = np.empty(self.nsteps)
        check(self._lib.vh_adjoint_info(self._h, C.byref(f), C.byref(b), C.byref(nbytes),
                                        dts.ctypes.data_as(C.POINTER(C.c_double))))
        return {"forward_ms": f.value, "backward_ms": b.value, "device_bytes": nbytes.value, "dts": dts}
