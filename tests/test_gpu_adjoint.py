"""GPU reverse mode (Float64) against the oracle's reverse mode and against
central finite differences.

The GPU adjoint gathers per cell where the oracle scatters per face, so the two
agree to round-off (not bit for bit); the forward states they differentiate are
bit-identical (exact profile). Finite differences are taken on states with
non-degenerate slopes (see test_adjoint_oracle.py for why states on minmod's
kink set are only directionally differentiable).
"""
import numpy as np
import pytest

from conftest import load_vespa
from helpers import prim2cons, random_state

V = load_vespa()
from vespa_jl_amd import adjoint as AD  # noqa: E402

pytestmark = pytest.mark.gpu

NAMES = {0: "periodic", 1: "outflow", 2: "reflecting"}


def _sim(dims, bcs, U):
    rank = len(dims)
    prob = V.Problem("adjoint", tuple(dims), tuple((0.0, 1.0) for _ in dims), 1.4, bcs,
                     lambda x, y, z: (1.0, 0.0, 0.0, 0.0, 1.0), 1.0, 0.4)
    sim = V.Simulation(V.UniformMesh(tuple(dims), prob.domain), prob)
    sim.upload(U.reshape((5,) + tuple(reversed(dims))))
    return sim


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = [
    ((40,), "outflow", 0),
    ((33,), "reflecting", 1),
    ((24, 20), "periodic", 2),
    ((17, 13), [("reflecting", "outflow"), ("outflow", "reflecting")], 3),
    ((8, 6, 5), [("periodic", "periodic"), ("reflecting", "outflow"), ("outflow", "reflecting")], 4),
    ((6, 7, 9), "periodic", 5),
]


@pytest.mark.parametrize("dims,bcs,seed", CASES)
def test_step_vjp_matches_oracle(O, dims, bcs, seed):
    print(f"step vjp {dims} {bcs}", flush=True)
    shape = (5,) + tuple(reversed(dims))
    U = random_state(shape[1:], seed=seed, vmax=0.4)
    g = O.make_grid(dims, [(0.0, 1.0)] * len(dims), bcs, 1.4)
    dt = 0.5 * O.compute_dt(g, U, 0.4)
    w = np.random.default_rng(seed).normal(size=shape)
    ref = O.step_vjp(g, U, dt, w)
    sim = _sim(dims, bcs, U)
    got = AD.step_vjp(sim, U, dt, w)
    assert _rel(got, ref) < 1e-12
    sim.close()


def _smooth_prim(n, seed):
    rng = np.random.default_rng(seed)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)

    def mode():
        k = rng.integers(1, 3, 2)
        return np.sin(2 * np.pi * (k[0] * X + k[1] * Y) + rng.uniform(0, 2 * np.pi))
    return np.stack([1.0 + 0.2 * mode(), 0.3 * mode(), 0.3 * mode(), 0.0 * X, 1.0 + 0.1 * mode()])


def test_rollout_gradient_matches_oracle_and_fd(O):
    """Checkpointed GPU rollout (24 steps, checkpoints every 5) against the
    oracle's chained VJPs over the same dt schedule, and against central
    differences of the oracle rollout on 6 random directions."""
    n, nst = 32, 24
    print("rollout gradient", flush=True)
    U0 = np.ascontiguousarray(prim2cons(_smooth_prim(n, 1), 1.4))
    target = _smooth_prim(n, 2)[0]
    sim = _sim((n, n), "periodic", U0)
    ro = AD.Rollout(sim, nst, every=5)
    UF = ro.forward()
    info = ro.info()
    dts = info["dts"]
    w = np.zeros_like(UF)
    w[0] = 2.0 * (UF[0] - target)
    G = ro.backward(w)
    # oracle on the same schedule: identical forward states, chained VJPs
    g = O.make_grid((n, n), [(0.0, 1.0)] * 2, "periodic", 1.4)
    states, U = [], U0.copy()
    for dt in dts:
        states.append(U.copy())
        O.step(g, U, dt)
    np.testing.assert_array_equal(U, UF)
    wr = w.copy()
    for k in reversed(range(nst)):
        wr = O.step_vjp(g, states[k], dts[k], wr)
    assert _rel(G, wr) < 1e-11

    def loss(x):
        y = np.ascontiguousarray(x.copy())
        for dt in dts:
            O.step(g, y, dt)
        return np.sum((y[0] - target) ** 2)
    rng = np.random.default_rng(3)
    for _ in range(6):
        v = rng.normal(size=U0.shape) * 0.01
        fd = (loss(U0 + 1e-6 * v) - loss(U0 - 1e-6 * v)) / 2e-6
        assert abs(np.sum(G * v) - fd) / abs(fd) < 1e-5
    after = ro.info()
    assert after["forward_ms"] > 0 and after["backward_ms"] > 0 and after["device_bytes"] > 0
    # the backward sweep is replayed from its captured graph: same gradient again,
    # and again after a fresh forward run from the same state
    np.testing.assert_array_equal(ro.backward(w), G)
    sim.upload(U0)
    np.testing.assert_array_equal(ro.forward(), UF)
    np.testing.assert_array_equal(ro.backward(w), G)
    ro.close()
    sim.close()


def test_rrule_step_pullback():
    n = 16
    U0 = np.ascontiguousarray(prim2cons(_smooth_prim(n, 4), 1.4))
    sim = _sim((n, n), "periodic", U0)
    dt = V.compute_dt(sim)
    U1, pb = AD.rrule_step(sim, dt)
    w = np.random.default_rng(0).normal(size=U1.shape)
    G = pb(w)
    assert G.shape == U0.shape and np.isfinite(G).all()
    np.testing.assert_array_equal(G, AD.step_vjp(sim, U0, dt, w))
    sim.close()


def test_float32_grid_is_rejected():
    n = 8
    U0 = np.ascontiguousarray(prim2cons(_smooth_prim(n, 4), 1.4)).astype(np.float32)
    prob = V.Problem("adjoint", (n, n), ((0.0, 1.0),) * 2, 1.4, "periodic",
                     lambda x, y, z: (1.0, 0.0, 0.0, 0.0, 1.0), 1.0, 0.4)
    sim = V.Simulation(V.UniformMesh((n, n), prob.domain), prob, eltype=np.float32)
    with pytest.raises(ValueError, match="Float64"):
        AD.Rollout(sim, 4)
    sim.close()


@pytest.mark.parametrize("dims", [(64, 32), (32, 16)])
def test_fused_tile_form_matches_three_pass_form_and_oracle(O, dims):
    """The single-kernel 2D periodic form against the three-pass form and the
    oracle (both forms differ from the oracle's scatter only in summation order)."""
    print(f"fused jat {dims}", flush=True)
    shape = (5,) + tuple(reversed(dims))
    U = random_state(shape[1:], seed=11, vmax=0.4)
    g = O.make_grid(dims, [(0.0, 1.0)] * 2, "periodic", 1.4)
    dt = 0.5 * O.compute_dt(g, U, 0.4)
    w = np.random.default_rng(12).normal(size=shape)
    ref = O.step_vjp(g, U, dt, w)
    sim = _sim(dims, "periodic", U)
    AD.set_form(sim, True)
    fused = AD.step_vjp(sim, U, dt, w)
    AD.set_form(sim, False)
    three = AD.step_vjp(sim, U, dt, w)
    assert _rel(fused, ref) < 1e-12
    assert _rel(three, ref) < 1e-12
    sim.close()
