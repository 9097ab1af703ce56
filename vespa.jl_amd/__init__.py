"""vespa.jl_amd — MI355X-native (gfx950) ghost-free flux-divergence hot path of Vespa.jl.

The product is ``libvespahip.so`` (hand-written HIP kernels behind the C ABI in
``include/vespahip.h``); this package is its host-side mirror of the reference's
flux-backend interface (see ``hydro.py``). The directory name contains a dot, so
load it with ``importlib`` under the module name ``vespa_jl_amd`` (as the tests,
``bench.py`` and ``__graft_entry__.py`` do).
"""
from ._lib import (DeviceBuffer, BC_OUTFLOW, BC_PARENT, BC_PERIODIC, BC_REFLECTING, EXPORTED, LIB_PATH, VH_EXACT, VH_F32, VH_F64, VH_FAST,
                   VespaHIPError, check, device_count, device_cu_count, load_library)
from .hydro import (BoundaryFluxRegister, HIPFlux, IdealHydro, Outflow, ParentGhost, Periodic, Problem, Reflecting,
                    Simulation, bflux_count, bflux_plane, parent_ghost_states, set_parent_ghost,
                    UniformMesh, compute_dt, compute_face_fluxes, cons2prim_array, conserved_totals,
                    dump_fields, evolve, prim2cons_array, run_steps, sod_problem, step)
from .gravity import GravityField, enable_gravity, solve_poisson
from .muscl import MusclGrid, cons_to_prim, muscl_step_3d, prim_to_cons, total_field

__all__ = [
    "DeviceBuffer",
    "BC_OUTFLOW", "BC_PARENT", "BC_PERIODIC", "BC_REFLECTING", "EXPORTED", "LIB_PATH", "VH_EXACT", "VH_F32", "VH_F64", "VH_FAST",
    "VespaHIPError", "check", "device_count", "device_cu_count", "load_library",
    "BoundaryFluxRegister", "bflux_count", "bflux_plane", "ParentGhost", "parent_ghost_states", "set_parent_ghost",
    "HIPFlux", "IdealHydro", "Outflow", "Periodic", "Problem", "Reflecting", "Simulation",
    "UniformMesh", "compute_dt", "compute_face_fluxes", "cons2prim_array", "conserved_totals",
    "dump_fields", "evolve", "prim2cons_array", "run_steps", "sod_problem", "step",
    "GravityField", "enable_gravity", "solve_poisson",
    "MusclGrid", "cons_to_prim", "muscl_step_3d", "prim_to_cons", "total_field",
]
