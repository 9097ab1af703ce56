"""Boundary-flux recording on the GPU (step!(sim, dt; bflux), src/reflux.jl:105-132):
bit-identical to the oracle's register in the exact profile, and the
conservation identity of test/test_boundary_flux.jl:9-36 (Δmass ≡ Σ_lo − Σ_hi)."""
import numpy as np
import pytest

from conftest import load_vespa
from helpers import random_state

V = load_vespa()

pytestmark = pytest.mark.gpu

CASES = [
    ((64,), "outflow"),
    ((20, 12), [("outflow", "reflecting"), ("reflecting", "outflow")]),
    ((12, 10, 9), [("outflow", "outflow"), ("periodic", "periodic"), ("reflecting", "outflow")]),
    ((40, 9, 7), "outflow"),
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dims,bcs", CASES)
def test_bflux_matches_oracle(O, dims, bcs, dtype):
    print(f"bflux {dims} {np.dtype(dtype).name}", flush=True)
    shape = tuple(reversed(dims))
    U = random_state(shape, seed=len(dims), vmax=0.4).astype(dtype)
    g = O.make_grid(dims, [(0.0, 1.0)] * len(dims), bcs, 1.4)
    prob = V.Problem("bflux", tuple(dims), tuple((0.0, 1.0) for _ in dims), 1.4, bcs,
                     lambda x, y, z: (1.0, 0.0, 0.0, 0.0, 1.0), 1.0, 0.3)
    sim = V.Simulation(V.UniformMesh(tuple(dims), prob.domain), prob, eltype=dtype)
    sim.upload(U)
    reg = V.BoundaryFluxRegister(sim)
    ref = np.zeros_like(reg.flux)
    assert ref.shape[0] == O.bface_count(g)
    Uo = np.ascontiguousarray(U.copy())
    for _ in range(3):
        dt = O.compute_dt(g, Uo, 0.3, floor64=dtype == np.float32)
        O.step(g, Uo, dt, floor64=dtype == np.float32, bflux=ref)
        V.step(sim, dt, bflux=reg)
    np.testing.assert_array_equal(sim.state(), Uo)
    np.testing.assert_array_equal(reg.flux, ref)
    sim.close()


def test_bflux_conservation_identity():
    """The reference's ramp problem: 64 outflow cells, ρ = 1 + x, v = 1, p = 2,
    20 steps of dt = 5e-4."""
    prob = V.Problem("ramp", (64,), ((0.0, 1.0),), 1.4, "outflow",
                     lambda x, y, z: (1.0 + x, 1.0, 0.0, 0.0, 2.0), 1.0, 0.3)
    sim = V.Simulation(V.UniformMesh((64,), prob.domain), prob)
    mass = lambda: float(sim.state()[0].sum() / 64)
    m0, lo, hi = mass(), 0.0, 0.0
    for _ in range(20):
        reg = V.BoundaryFluxRegister(sim)
        V.step(sim, 5e-4, bflux=reg)
        lo += reg.plane(1, "lo")[0]
        hi += reg.plane(1, "hi")[0]
    dm = mass() - m0
    assert abs(dm) > 1e-3
    assert abs(dm - (lo - hi)) < 1e-12
    sim.close()
