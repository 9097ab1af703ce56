"""The adjoint-rollout workload on the GPU: a 2D periodic Kelvin–Helmholtz shear
layer (γ = 1.4, vx = ±0.5 with tanh width 0.025, densities 1 and 2, p = 2.5,
seeded vy modes |k| ≤ 4 of amplitude 0.01), SSP-RK2 steps at CFL 0.4 with the
device dt, the loss L = Σ (ρ_final − ρ_target)² against the final density of a
run from a second seeded state, and the gradient w.r.t. the initial conserved
fields from the checkpointed hand-written adjoint.

Checks:
* 128² × 200 steps, checkpoints every 20: the forward states equal the oracle's
  bit for bit, and the GPU gradient equals the oracle's chained step VJPs on the
  same dt schedule (≤ 1e-10 of its largest entry).
* The transpose identity ⟨G, v⟩ = ⟨w, J v⟩ for 16 random directions v, with J v
  from the oracle's forward mode (dual numbers, the same branch rules) on the same
  schedule — 128² × 200 steps and a 20-step segment at the full 512². Finite
  differences cannot pin the gradient on this state (its slopes sit on minmod's
  kink and its faces on Davis min/max ties), the identity can (≤ 1e-10).
* 512² × 200 steps, checkpoints every 20: the backward pass costs at most 3× the
  forward pass (HIP event times of the library's own rollout).
"""
import sys

import numpy as np
import pytest

from conftest import load_vespa

V = load_vespa()
from vespa_jl_amd import adjoint as AD  # noqa: E402
from vespa_jl_amd import problems as P  # noqa: E402

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sim(n, U):
    prob = V.Problem("kh", (n, n), ((0.0, 1.0),) * 2, 1.4, "periodic",
                     lambda x, y, z: (1.0, 0.0, 0.0, 0.0, 1.0), 1.0, 0.4)
    sim = V.Simulation(V.UniformMesh((n, n), prob.domain), prob)
    sim.set_mode("exact")
    sim.upload(U)
    return sim


def _target(n, steps):
    s = _sim(n, np.ascontiguousarray(P.kh_state(n, seed=8)))
    V.run_steps(s, steps)
    rho = s.state()[0].copy()
    s.close()
    return rho


def _gpu_rollout(n, steps, every=20):
    """Returns U0, final state, w = ∂L/∂U_final, gradient, dt schedule, info."""
    U0 = np.ascontiguousarray(P.kh_state(n, seed=7))
    target = _target(n, steps)
    sim = _sim(n, U0)
    ro = AD.Rollout(sim, steps, every=every)
    UF = ro.forward()
    w = np.zeros_like(UF)
    w[0] = 2.0 * (UF[0] - target)
    G = ro.backward(w)
    info = ro.info()
    ro.close()
    sim.close()
    return U0, UF, w, G, info["dts"], info


def _oracle_chain(O, g, U0, dts, w):
    states, U = [], U0.copy()
    for dt in dts:
        states.append(U.copy())
        O.step(g, U, float(dt))
    G = w.copy()
    for k in reversed(range(len(dts))):
        G = O.step_vjp(g, states[k], float(dts[k]), G)
    return U, G


def _transpose_gap(O, g, U0, dts, w, G, UF, seed):
    rng = np.random.default_rng(seed)
    nd = O.jvp_max_directions()
    Vd = rng.normal(size=(nd,) + U0.shape)
    Un, JV = O.rollout_jvp(g, U0, dts, Vd, threads=16)
    assert np.array_equal(Un, UF), "forward mode's states differ from the GPU rollout's"
    lhs = Vd.reshape(nd, -1) @ G.ravel()
    rhs = JV.reshape(nd, -1) @ w.ravel()
    return float(np.abs(lhs - rhs).max() / np.abs(rhs).max())


def test_kh128_rollout_gradient_matches_oracle_chain_and_forward_mode(O):
    n, steps = 128, 200
    _log(f"KH {n}^2 x {steps} steps, checkpoints every 20")
    U0, UF, w, G, dts, _ = _gpu_rollout(n, steps)
    g = O.make_grid((n, n), [(0.0, 1.0)] * 2, "periodic", 1.4)
    Uo, Go = _oracle_chain(O, g, U0, dts, w)
    assert np.array_equal(Uo, UF)
    rel = float(np.abs(G - Go).max() / np.abs(Go).max())
    _log(f"  gradient vs oracle chain: {rel:.3e}")
    assert rel <= 1e-10
    gap = _transpose_gap(O, g, U0, dts, w, G, UF, seed=11)
    _log(f"  <G,v> vs <w,Jv> over 16 directions: {gap:.3e}")
    assert gap <= 1e-10


def test_kh512_segment_gradient_matches_oracle_and_forward_mode(O):
    n, steps = 512, 20
    _log(f"KH {n}^2 x {steps} steps (one checkpoint segment)")
    U0, UF, w, G, dts, _ = _gpu_rollout(n, steps)
    g = O.make_grid((n, n), [(0.0, 1.0)] * 2, "periodic", 1.4)
    Uo, Go = _oracle_chain(O, g, U0, dts, w)
    assert np.array_equal(Uo, UF)
    rel = float(np.abs(G - Go).max() / np.abs(Go).max())
    _log(f"  gradient vs oracle chain: {rel:.3e}")
    assert rel <= 1e-10
    gap = _transpose_gap(O, g, U0, dts, w, G, UF, seed=12)
    _log(f"  <G,v> vs <w,Jv> over 16 directions: {gap:.3e}")
    assert gap <= 1e-10


def test_kh512_full_rollout_backward_within_three_forwards():
    n, steps = 512, 200
    _log(f"KH {n}^2 x {steps} steps, checkpoints every 20: timing")
    U0 = np.ascontiguousarray(P.kh_state(n, seed=7))
    target = _target(n, steps)
    sim = _sim(n, U0)
    ro = AD.Rollout(sim, steps, every=20)
    ratios = []
    for _ in range(3):                      # the first pass warms up and captures the graph
        sim.upload(U0)
        UF = ro.forward()
        w = np.zeros_like(UF)
        w[0] = 2.0 * (UF[0] - target)
        G = ro.backward(w)
        info = ro.info()
        ratios.append(info["backward_ms"] / info["forward_ms"])
    ro.close()
    sim.close()
    _log(f"  backward/forward ratios {ratios}, device bytes {info['device_bytes']}")
    assert np.isfinite(G).all() and np.abs(G).max() > 0
    assert min(ratios[1:]) <= 3.0
    # rollout memory (10 checkpoints, one segment's stage states, the last segment
    # kept from the forward pass, reverse-mode work) well below storing both
    # stage states of all 200 steps
    assert info["device_bytes"] < 0.3 * steps * 2 * U0.nbytes
