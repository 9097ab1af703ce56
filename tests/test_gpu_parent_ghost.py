"""ParentGhost boundaries on the GPU (lib/MeshInterface/src/MeshInterface.jl:86-99,
src/driver.jl:127-134): the ghost states of a ParentGhost side, given per boundary
face through vh_grid_set_parent_ghost (or evaluated from a Python closure), are
the outer neighbour of the boundary cells' slopes and the outer state of their
faces. Against the oracle (pinned in test_parent_ghost_oracle.py) bit for bit in
the exact profile: states over three steps in 1D/2D/3D with ParentGhost sides
mixed with outflow, reflecting and periodic ones, Float64 and Float32; the
boundary-flux register; the gravity form; the closure path; the refusals."""
import numpy as np
import pytest

from conftest import load_vespa
from helpers import random_state

V = load_vespa()

pytestmark = pytest.mark.gpu
G14 = 1.4

CASES = [
    ((64,), [("parent", "outflow")]),
    ((20, 12), [("outflow", "parent"), ("parent", "reflecting")]),
    ((12, 10, 9), [("parent", "parent"), ("periodic", "periodic"), ("reflecting", "parent")]),
    ((40, 9, 7), [("parent", "outflow"), ("outflow", "parent"), ("parent", "outflow")]),
]


def _lib_bcs(bcs):
    return [tuple(V.ParentGhost(lambda a, s, c: None) if b == "parent" else b for b in p) for p in bcs]


def _states(O, g, dtype, seed):
    """Smooth-ish primitive states for every boundary face (ρ, p > 0)."""
    rng = np.random.default_rng(seed)
    n = O.bface_count(g)
    st = np.empty((n, 5))
    st[:, 0] = rng.uniform(0.6, 1.4, n)
    st[:, 1:4] = rng.normal(0.0, 0.3, (n, 3))
    st[:, 4] = rng.uniform(0.5, 1.5, n)
    return st.astype(dtype)


def _sim(dims, bcs, dtype):
    prob = V.Problem("pg", tuple(dims), tuple((0.0, 1.0) for _ in dims), G14, _lib_bcs(bcs),
                     lambda x, y, z: (1.0, 0.0, 0.0, 0.0, 1.0), 1.0, 0.3)
    sim = V.Simulation(V.UniformMesh(tuple(dims), prob.domain), prob, eltype=dtype)
    sim._parent = {}      # states are handed over directly below
    return sim


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dims,bcs", CASES)
def test_parent_ghost_steps_and_register_match_oracle(O, dims, bcs, dtype):
    print(f"parent ghost {dims} {bcs} {np.dtype(dtype).name}", flush=True)
    f64 = dtype == np.float32
    shape = tuple(reversed(dims))
    U = random_state(shape, seed=len(dims) + 3, vmax=0.4).astype(dtype)
    g = O.make_grid(dims, [(0.0, 1.0)] * len(dims), bcs, G14)
    st = _states(O, g, dtype, seed=len(dims))
    O.set_parent_ghost(g, st)
    sim = _sim(dims, bcs, dtype)
    try:
        V.set_parent_ghost(sim, st)
        sim.upload(U)
        Uo = np.ascontiguousarray(U.copy())
        for _ in range(3):
            dt = O.compute_dt(g, Uo, 0.3, floor64=f64)
            O.step(g, Uo, dt, floor64=f64)
            V.step(sim, dt)
        np.testing.assert_array_equal(sim.state(), Uo)
        # the boundary-flux register with ParentGhost faces
        reg = V.BoundaryFluxRegister(sim)
        ref = np.zeros_like(reg.flux)
        dt = O.compute_dt(g, Uo, 0.3, floor64=f64)
        O.step(g, Uo, dt, floor64=f64, bflux=ref)
        V.step(sim, dt, bflux=reg)
        np.testing.assert_array_equal(sim.state(), Uo)
        np.testing.assert_array_equal(reg.flux, ref)
    finally:
        sim.close()


def test_parent_ghost_with_gravity_matches_oracle(O):
    dims, bcs = (16, 12, 10), [("parent", "outflow"), ("periodic", "periodic"), ("outflow", "parent")]
    print("parent ghost with gravity", flush=True)
    shape = tuple(reversed(dims))
    U = random_state(shape, seed=5, vmax=0.3)
    g = O.make_grid(dims, [(0.0, 1.0)] * 3, bcs, G14)
    st = _states(O, g, np.float64, seed=9)
    O.set_parent_ghost(g, st)
    sim = _sim(dims, bcs, np.float64)
    try:
        V.set_parent_ghost(sim, st)
        sim.upload(U)
        field = V.enable_gravity(sim, G=1.0, bcs=["neumann", "periodic", "neumann"], maxiter=400, tol=1e-10)
        gper = 0b010
        phi = np.random.default_rng(2).normal(size=shape) * 0.01
        field.set_phi(phi)
        dt = 0.5 * O.compute_dt_grav(g, np.ascontiguousarray(U), 0.3, 1.0)
        Uo = np.ascontiguousarray(U.copy())
        for _ in range(2):
            O.step_grav(g, Uo, dt, gper, phi.ravel())
            V.step(sim, dt)
        np.testing.assert_array_equal(sim.state(), Uo)
        field.close()
    finally:
        sim.close()


def test_parent_ghost_closure_path(O):
    """ParentGhost(ghost) in the Problem's bcs: the Simulation evaluates the
    closure (1-based axis and cell, "lo"/"hi") for every face before the step."""
    dims = (10, 8)
    bcs = [("parent", "outflow"), ("periodic", "periodic")]
    calls = []

    def ghost(axis, side, cell):
        calls.append((axis, side, cell))
        return (1.0 + 0.05 * cell[1], 0.3, -0.1 * cell[1], 0.0, 0.8)

    prob = V.Problem("pg", dims, ((0.0, 1.0), (0.0, 1.0)), G14, [(V.ParentGhost(ghost), "outflow"),
                                                                ("periodic", "periodic")],
                     lambda x, y, z: (1.0, 0.0, 0.0, 0.0, 1.0), 1.0, 0.3)
    sim = V.Simulation(V.UniformMesh(dims, prob.domain), prob)
    try:
        U = random_state(tuple(reversed(dims)), seed=1, vmax=0.3)
        sim.upload(U)
        V.step(sim, 0.01)
        assert sorted(set(calls)) == [(1, "lo", (1, j)) for j in range(1, 9)]
        g = O.make_grid(dims, [(0.0, 1.0)] * 2, bcs, G14)
        st = np.zeros((O.bface_count(g), 5))
        for j in range(8):
            st[j] = ghost(1, "lo", (1, j + 1))
        O.set_parent_ghost(g, st)
        Uo = np.ascontiguousarray(U.copy())
        O.step(g, Uo, 0.01)
        np.testing.assert_array_equal(sim.state(), Uo)
    finally:
        sim.close()


def test_parent_ghost_refusals():
    dims = (12, 6)
    sim = _sim(dims, [("parent", "outflow"), ("periodic", "periodic")], np.float64)
    try:
        with pytest.raises(V.VespaHIPError, match="ParentGhost"):
            V.step(sim, 0.01)           # no states yet
        V.set_parent_ghost(sim, np.ones((V.bflux_count(sim), 5)))
        V.step(sim, 0.01)
        from vespa_jl_amd import adjoint as A
        with pytest.raises(V.VespaHIPError, match="ParentGhost"):
            A.step_vjp(sim, sim.state(), 0.01, np.ones((5,) + tuple(reversed(dims))))
    finally:
        sim.close()
    plain = _sim(dims, [("outflow", "outflow"), ("periodic", "periodic")], np.float64)
    try:
        with pytest.raises(V.VespaHIPError, match="no side"):
            V.set_parent_ghost(plain, np.ones((V.bflux_count(plain), 5)))
    finally:
        plain.close()
