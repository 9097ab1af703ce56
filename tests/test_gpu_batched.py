"""The tier-1 seam's whole-run gate (test/gpu/test_ka_flux_gpu.jl:93-102): a
simulation whose accumulate_flux! runs through `sim.flux = HIPFlux()` (host
gather → GPU HLLC batch → host scatter) against the fused device path on the
same problem. Float64: bit for bit over Sod-128 to t = 0.2 and over a few steps
of 2D/3D grids with every boundary kind (the reference only asks for 1e-9).
Float32: the fused kernels evaluate the pressure in double (the FP32 form's
floor policy) where the batched gather evaluates it in Float32, so the two agree
to Float32 round-off."""
import sys

import numpy as np
import pytest

from conftest import load_vespa
from helpers import random_state

V = load_vespa()

pytestmark = pytest.mark.gpu
NAMES = {0: "periodic", 1: "outflow", 2: "reflecting"}


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_sod128_batched_seam_equals_fused_path():
    _log("Sod-128 to t=0.2: sim.flux = HIPFlux() vs fused")
    out = []
    for flux in (V.HIPFlux(), None):
        prob = V.sod_problem(128)
        sim = V.Simulation(V.UniformMesh(prob.dims, prob.domain), prob, flux=flux)
        V.evolve(sim)
        out.append((sim.state(), sim.t, sim.step_count))
        sim.close()
    (Ub, tb, kb), (Uf, tf, kf) = out
    assert kb == kf and tb == tf
    assert np.array_equal(Ub, Uf)


@pytest.mark.parametrize("dims,bcs", [((40, 24), [(1, 2), (0, 0)]),
                                       ((12, 10, 8), [(0, 0), (2, 1), (1, 2)])])
def test_batched_seam_equals_fused_path_nd(dims, bcs):
    _log(f"batched seam {dims} {bcs}")
    U = random_state(tuple(reversed(dims)), seed=3, vmax=0.4)
    out = []
    for flux in (V.HIPFlux(), None):
        prob = V.Problem("seam", dims, tuple((0.0, 1.0) for _ in dims), 1.4,
                         [(NAMES[a], NAMES[b]) for a, b in bcs], lambda x, y, z: (1.0, 0.0, 0.0, 0.0, 1.0), 1.0, 0.4)
        sim = V.Simulation(V.UniformMesh(dims, prob.domain), prob, flux=flux)
        sim.upload(U)
        V.run_steps(sim, 4)
        out.append(sim.state())
        sim.close()
    assert np.array_equal(out[0], out[1])


def test_batched_seam_float32_within_round_off():
    dims = (32, 20)
    U = random_state((20, 32), seed=4, vmax=0.4, dtype=np.float32)
    out = []
    for flux in (V.HIPFlux(), None):
        prob = V.Problem("seam32", dims, ((0.0, 1.0),) * 2, 1.4, "outflow",
                         lambda x, y, z: (1.0, 0.0, 0.0, 0.0, 1.0), 1.0, 0.4)
        sim = V.Simulation(V.UniformMesh(dims, prob.domain), prob, eltype=np.float32, flux=flux)
        sim.upload(U)
        V.run_steps(sim, 3)
        out.append(sim.state().astype(np.float64))
        sim.close()
    d = np.abs(out[0] - out[1]).max(axis=(1, 2)) / np.abs(out[1]).max(axis=(1, 2))
    assert d.max() <= 1e-5
