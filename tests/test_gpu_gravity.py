"""Self-gravity on the GPU against the oracle (src/gravity.jl): the Laplacian
matrix-vector product, g = −∇φ, a step with the gravity source and compute_dt
with the free-fall limiter are bit-identical in the exact profile (FP64 and the
FP32 form); the CG solve agrees to its tolerance; the reference's physics
oracles (Fourier mode at second order, Jeans growth) hold on the device path."""
import numpy as np
import pytest

from conftest import load_vespa
from helpers import random_state

V = load_vespa()

pytestmark = pytest.mark.gpu

G53 = 5.0 / 3.0

CASES = [
    ((40,), "periodic", "periodic"),
    ((24, 20), "periodic", ["periodic", "neumann"]),
    ((12, 10, 8), "periodic", "periodic"),
    ((12, 10, 8), "outflow", "neumann"),
    ((16, 12, 6), [("periodic", "periodic"), ("outflow", "reflecting"), ("periodic", "periodic")],
     ["periodic", "neumann", "periodic"]),
]


def _sim(dims, bcs, U, dtype):
    prob = V.Problem("grav", tuple(dims), tuple((0.0, 1.0) for _ in dims), G53, bcs,
                     lambda x, y, z: (1.0, 0.0, 0.0, 0.0, 1.0), 1.0, 0.3)
    sim = V.Simulation(V.UniformMesh(tuple(dims), prob.domain), prob, eltype=dtype)
    sim.upload(U)
    return sim


def _gper(rank, pbcs):
    pb = [pbcs] * rank if isinstance(pbcs, str) else pbcs
    return sum(1 << d for d in range(rank) if pb[d] == "periodic")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dims,bcs,pbcs", CASES)
def test_operator_accel_step_dt_bitwise(O, dims, bcs, pbcs, dtype):
    print(f"gravity parity {dims} {bcs} {pbcs} {np.dtype(dtype).name}", flush=True)
    shape = tuple(reversed(dims))
    U = random_state(shape, seed=len(dims) + 3, gamma=G53, vmax=0.3).astype(dtype)
    sim = _sim(dims, bcs, U, dtype)
    grav = V.enable_gravity(sim, G=2.0, bcs=pbcs)
    g = O.make_grid(dims, [(0.0, 1.0)] * len(dims), bcs, G53)
    gper = _gper(len(dims), pbcs)
    rng = np.random.default_rng(5)
    x = rng.normal(size=shape).astype(dtype)
    np.testing.assert_array_equal(grav.apply_laplacian(x).ravel(), O.poisson_apply(g, gper, x.ravel()))
    phi = (0.05 * rng.normal(size=shape)).astype(dtype)
    grav.set_phi(phi)
    np.testing.assert_array_equal(grav.accel().reshape(3, -1), O.grav_accel(g, gper, phi.ravel()))
    f64 = dtype == np.float32
    dt_ref = O.compute_dt_grav(g, U, 0.3, 2.0, floor64=f64)
    assert V.compute_dt(sim) == dt_ref
    Uo = np.ascontiguousarray(U.copy())
    for _ in range(2):
        O.step_grav(g, Uo, dt_ref, gper, phi.ravel(), floor64=f64)
        V.step(sim, dt_ref)
    np.testing.assert_array_equal(sim.state(), Uo)
    grav.close()
    sim.close()


@pytest.mark.parametrize("dims,pbcs", [((64,), "periodic"), ((16, 12, 10), "periodic"), ((20, 16), "neumann")])
def test_cg_solve_matches_oracle(O, dims, pbcs):
    print(f"gravity cg {dims} {pbcs}", flush=True)
    shape = tuple(reversed(dims))
    U = random_state(shape, seed=11, gamma=G53, vmax=0.2)
    sim = _sim(dims, "periodic", U, np.float64)
    # all-Neumann φ is singular too: the mean projection makes its right-hand side consistent
    grav = V.enable_gravity(sim, G=1.0, bcs=pbcs, tol=1e-10, maxiter=2000, project_mean=True)
    it, rr = grav.solve()
    g = O.make_grid(dims, [(0.0, 1.0)] * len(dims), "periodic", G53)
    phi_o = np.zeros(int(np.prod(dims)))
    it_o, rr_o = O.poisson_solve(g, _gper(len(dims), pbcs), U, 1.0, True, 2000, 1e-10, phi_o)
    assert rr <= 1e-10 and rr_o <= 1e-10
    assert abs(it - it_o) <= max(2, it_o // 20)
    phi = grav.phi().ravel()
    assert np.abs(phi - phi_o).max() <= 1e-8 * np.abs(phi_o).max()
    # warm start: a second solve of the same density converges at once
    it2, rr2 = grav.solve()
    assert it2 <= 2 and rr2 <= 1e-10
    sim.close()


def _fourier_gpu(n):
    x = (np.arange(n) + 0.5) / n
    k = 2 * np.pi
    W = np.stack([1.0 + 0.1 * np.cos(k * x), 0 * x, 0 * x, 0 * x, 1.0 + 0 * x])
    U = V.prim2cons_array(W, G53, np.float64)
    sim = _sim((n,), "periodic", U, np.float64)
    grav = V.enable_gravity(sim, tol=1e-12, maxiter=5000)
    _, rr = grav.solve()
    err = np.abs(grav.phi() - (-4 * np.pi * 0.1 / k ** 2) * np.cos(k * x)).mean()
    sim.close()
    return err, rr


def test_fourier_mode_second_order():
    """test/test_gravity_poisson.jl:54-80 on the device solver."""
    e64, rr = _fourier_gpu(64)
    e128, _ = _fourier_gpu(128)
    assert rr < 1e-10 and e64 < 5e-3 and e64 / e128 > 3.5


def test_jeans_growth_rate():
    """test/test_gravity_hydro.jl:54-87 through the device run loop (Poisson
    solve, free-fall dt and the sourced step every step)."""
    n, p0, A = 256, 0.06, 1e-4
    k = 2 * np.pi
    x = (np.arange(n) + 0.5) / n
    W = np.stack([1.0 + A * np.cos(k * x), 0 * x, 0 * x, 0 * x, p0 + 0 * x])
    sim = _sim((n,), "periodic", V.prim2cons_array(W, G53, np.float64), np.float64)
    V.enable_gravity(sim)
    s_exact = np.sqrt(4 * np.pi - G53 * p0 * k ** 2)
    tmax = 2.0 / s_exact
    ts, amps = [], []
    ck = np.cos(k * x)
    while sim.t < tmax * (1 - 1e-12):
        V.run_steps(sim, 1)
        ts.append(sim.t)
        amps.append(((sim.state()[0] - 1.0) * ck).sum() / (ck * ck).sum())
    h = len(ts) // 2
    s_meas = np.polyfit(ts[h:], np.log(np.abs(amps[h:]) / A), 1)[0]
    assert abs(s_meas - s_exact) / s_exact < 0.08
    sim.close()


def test_disable_and_reverse_mode_guard(O):
    dims = (16, 12)
    U = random_state((12, 16), seed=2, gamma=G53, vmax=0.3)
    sim = _sim(dims, "periodic", U, np.float64)
    grav = V.enable_gravity(sim)
    from vespa_jl_amd import adjoint as AD
    with pytest.raises(V.VespaHIPError, match="self-gravity"):
        AD.step_vjp(sim, U, 1e-3, np.ones_like(U))
    grav.close()
    g = O.make_grid(dims, [(0.0, 1.0)] * 2, "periodic", G53)
    dt = O.compute_dt(g, U, 0.3)
    assert V.compute_dt(sim) == dt
    Uo = np.ascontiguousarray(U.copy())
    O.step(g, Uo, dt)
    V.step(sim, dt)
    np.testing.assert_array_equal(sim.state(), Uo)
    sim.close()
