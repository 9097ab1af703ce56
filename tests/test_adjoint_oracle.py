"""Pins of the oracle's reverse mode (the derivative of its own SSP-RK2 step with
dt fixed) by central finite differences of the forward oracle.

Every piece is checked in every branch the forward takes: cons2prim with and
without the pressure floor; HLLC in its four regimes (supersonic left/right,
star state of either side) along each axis; the accumulate operator and the
whole step on small grids with periodic, outflow and reflecting boundaries;
a multi-step rollout with a fixed dt schedule. A dropped term, a wrong sign or
a transposed index shows up as an O(1) relative error; the finite differences
agree to ~1e-8.
"""
import numpy as np
import pytest

from helpers import random_state, smooth_state

EPS = 1e-6


def fd_dir(f, x, v, eps=EPS):
    return (f(x + eps * v) - f(x - eps * v)) / (2 * eps)


def rel_err(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def test_cons2prim_vjp_matches_fd(O):
    rng = np.random.default_rng(0)
    g = 1.4
    for floor in (False, True):
        for _ in range(20):
            W = np.array([rng.uniform(0.5, 2), *rng.uniform(-1, 1, 3), rng.uniform(0.5, 2)])
            U = O.prim2cons(W, g)
            if floor:       # energy below the kinetic part: the floor is the larger argument
                kin = 0.5 * W[0] * np.sum(W[1:4] ** 2)
                U[4] = 0.5 * kin
            wb = rng.normal(size=5)
            ub = O.cons2prim_vjp(U, g, wb)
            for i in range(5):
                e = np.zeros(5)
                e[i] = 1.0
                h = EPS * max(abs(U[i]), 1e-3)
                fd = (wb @ O.cons2prim(U + h * e, g) - wb @ O.cons2prim(U - h * e, g)) / (2 * h)
                assert abs(ub[i] - fd) <= 1e-6 * max(1.0, abs(fd)), (floor, i, ub[i], fd)


def _regime(WL, WR, g, axis):
    cL = np.sqrt(g * WL[4] / WL[0])
    cR = np.sqrt(g * WR[4] / WR[0])
    uL, uR = WL[axis], WR[axis]
    SL, SR = min(uL - cL, uR - cR), max(uL + cL, uR + cR)
    if SL >= 0:
        return "supL"
    if SR <= 0:
        return "supR"
    Ss = ((WR[4] - WL[4]) + WL[0] * uL * (SL - uL) - WR[0] * uR * (SR - uR)) / (
        WL[0] * (SL - uL) - WR[0] * (SR - uR))
    return "starL" if Ss >= 0 else "starR"


@pytest.mark.parametrize("axis", [1, 2, 3])
def test_hllc_vjp_matches_fd_in_every_regime(O, axis):
    g = 1.4
    rng = np.random.default_rng(axis)
    seen = set()
    cases = []
    for shift in (3.0, -3.0, 0.3, -0.3):
        for _ in range(6):
            WL = np.array([rng.uniform(0.5, 2), *rng.uniform(-0.5, 0.5, 3), rng.uniform(0.5, 2)])
            WR = np.array([rng.uniform(0.5, 2), *rng.uniform(-0.5, 0.5, 3), rng.uniform(0.5, 2)])
            WL[axis] += shift
            WR[axis] += shift
            cases.append((WL, WR))
    for WL, WR in cases:
        reg = _regime(WL, WR, g, axis)
        seen.add(reg)
        fb = rng.normal(size=5)
        a, b = O.hllc_vjp(WL, WR, g, axis, fb)
        for side, W, grad in (("L", WL, a), ("R", WR, b)):
            for i in range(5):
                e = np.zeros(5)
                e[i] = 1.0

                def f(x):
                    return fb @ (O.hllc_flux(x, WR, g, axis) if side == "L" else O.hllc_flux(WL, x, g, axis))
                fd = fd_dir(f, W, e)
                assert abs(grad[i] - fd) <= 1e-6 * max(1.0, abs(fd)), (reg, side, i, grad[i], fd)
    assert seen == {"supL", "supR", "starL", "starR"}


GRIDS = [
    ((10,), "outflow", 0),
    ((9,), "reflecting", 1),
    ((6, 5), "periodic", 2),
    ((5, 4), "reflecting", 3),
    ((4, 3, 3), [("outflow", "reflecting"), ("periodic", "periodic"), ("reflecting", "outflow")], 4),
]


@pytest.mark.parametrize("dims,bcs,seed", GRIDS)
def test_accumulate_vjp_matches_fd(O, dims, bcs, seed):
    g = O.make_grid(dims, [(0.0, 1.0)] * len(dims), bcs, 1.4)
    shape = (5,) + tuple(reversed(dims))
    U = random_state(shape[1:], seed=seed, vmax=0.4)
    rng = np.random.default_rng(seed + 10)
    lam = rng.normal(size=shape)
    J = O.accumulate_vjp(g, U, lam)
    for _ in range(6):
        v = rng.normal(size=shape) * np.abs(U) * 0.1
        fd = fd_dir(lambda x: np.sum(lam * O.accumulate_flux(g, x)), U, v)
        assert rel_err(np.sum(J * v), fd) < 1e-6


@pytest.mark.parametrize("dims,bcs,seed", GRIDS)
def test_step_vjp_matches_fd(O, dims, bcs, seed):
    g = O.make_grid(dims, [(0.0, 1.0)] * len(dims), bcs, 1.4)
    shape = (5,) + tuple(reversed(dims))
    U = random_state(shape[1:], seed=seed + 20, vmax=0.4)
    dt = 0.5 * O.compute_dt(g, U, 0.4)
    rng = np.random.default_rng(seed + 30)
    w = rng.normal(size=shape)
    G = O.step_vjp(g, U, dt, w)

    def L(x):
        y = np.ascontiguousarray(x.copy())
        O.step(g, y, dt)
        return np.sum(w * y)
    for _ in range(6):
        v = rng.normal(size=shape) * np.abs(U) * 0.1
        assert rel_err(np.sum(G * v), fd_dir(L, U, v)) < 1e-6


def _smooth_prim(dims, seed, gamma, p_amp):
    """Smooth periodic primitive fields on the unit square (x fastest)."""
    rng = np.random.default_rng(seed)
    x = (np.arange(dims[0]) + 0.5) / dims[0]
    y = (np.arange(dims[1]) + 0.5) / dims[1]
    X, Y = np.meshgrid(x, y)

    def mode():
        k = rng.integers(1, 3, 2)
        return np.sin(2 * np.pi * (k[0] * X + k[1] * Y) + rng.uniform(0, 2 * np.pi))
    return np.stack([1.0 + 0.2 * mode(), 0.3 * mode(), 0.3 * mode(), 0.0 * X, 1.0 + p_amp * mode()])


def _rollout(O, g, U0, target, nsteps):
    dts, states = [], []
    U = U0.copy()
    for _ in range(nsteps):
        dt = O.compute_dt(g, U, 0.4)
        states.append(U.copy())
        dts.append(dt)
        O.step(g, U, dt)
    w = np.zeros_like(U)
    w[0] = 2.0 * (U[0] - target)
    for k in reversed(range(nsteps)):
        w = O.step_vjp(g, states[k], dts[k], w)

    def loss(x):
        y = np.ascontiguousarray(x.copy())
        for dt in dts:
            O.step(g, y, dt)
        return np.sum((y[0] - target) ** 2)
    return w, loss


def test_rollout_gradient_with_fixed_dt_schedule(O):
    """Chained step VJPs over 6 steps of a smooth periodic 2D state (the adjoint
    rollout at small size) against central differences of the rollout."""
    dims = (12, 10)
    g = O.make_grid(dims, [(0.0, 1.0)] * 2, "periodic", 1.4)
    U0 = O_prim2cons(_smooth_prim(dims, 5, 1.4, 0.1))
    target = _smooth_prim(dims, 6, 1.4, 0.1)[0]
    w, loss = _rollout(O, g, U0, target, 6)
    rng = np.random.default_rng(7)
    for _ in range(8):
        v = rng.normal(size=U0.shape) * 0.01
        assert rel_err(np.sum(w * v), fd_dir(loss, U0, v)) < 1e-5


def test_rollout_gradient_at_uniform_pressure_along_pressure_preserving_directions(O):
    """With the pressure exactly uniform (as in the shear-layer rollout) every
    pressure slope sits at minmod's kink, where the step is not differentiable
    along directions that perturb the pressure; along directions that keep it
    uniform the derivative exists and the adjoint (zero subgradient at the kink)
    must match central differences taken in primitive space."""
    dims = (12, 10)
    g = O.make_grid(dims, [(0.0, 1.0)] * 2, "periodic", 1.4)
    W0 = _smooth_prim(dims, 8, 1.4, 0.0)
    U0 = O_prim2cons(W0)
    target = _smooth_prim(dims, 9, 1.4, 0.0)[0]
    w, loss = _rollout(O, g, U0, target, 6)
    rng = np.random.default_rng(11)
    for _ in range(8):
        dW = rng.normal(size=W0.shape) * 0.01
        dW[4] = 0.0
        r, vx, vy, vz, p = W0
        dr, dvx, dvy, dvz, _ = dW
        # conserved direction = Jacobian of prim2cons at W0 applied to dW (dp = 0)
        dU = np.stack([dr, dr * vx + r * dvx, dr * vy + r * dvy, dr * vz + r * dvz,
                       0.5 * dr * (vx * vx + vy * vy + vz * vz) + r * (vx * dvx + vy * dvy + vz * dvz)])
        fd = (loss(O_prim2cons(W0 + EPS * dW)) - loss(O_prim2cons(W0 - EPS * dW))) / (2 * EPS)
        assert rel_err(np.sum(w * dU), fd) < 1e-5


def O_prim2cons(W, gamma=1.4):
    from helpers import prim2cons
    return np.ascontiguousarray(prim2cons(W, gamma))
