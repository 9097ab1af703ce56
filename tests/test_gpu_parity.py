"""GPU parity tests: the HIP library (through its C ABI) against the CPU oracle.

Bar: bit-exact for FP64 and for the FP32 form (whose oracle evaluates the
pressure floor in double, as the kernels do); documented tolerances only for the
order-dependent Float64 totals.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from helpers import random_faces, random_state, smooth_state

pytestmark = pytest.mark.gpu

NAMES = {0: "periodic", 1: "outflow", 2: "reflecting"}


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make_sim(V, dims, bcs, gamma, dtype, U=None):
    """A Simulation on [0,1]^rank with per-axis (lo, hi) BC codes and state U."""
    rank = len(dims)
    prob = V.Problem("parity", tuple(dims), tuple((0.0, 1.0) for _ in dims), gamma,
                     [(NAMES[a], NAMES[b]) for a, b in bcs],
                     lambda x, y, z: (1.0, 0.0, 0.0, 0.0, 1.0), 1.0, 0.4)
    sim = V.Simulation(V.UniformMesh(tuple(dims), prob.domain), prob, eltype=dtype)
    if U is not None:
        sim.upload(U)
    return sim


def oracle_grid(O, dims, bcs, gamma):
    return O.make_grid(dims, [(0.0, 1.0)] * len(dims), [(NAMES[a], NAMES[b]) for a, b in bcs], gamma)


# --------------------------------------------------------------------------
# tier 1: compute_face_fluxes!

@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nf", [1, 255, 4097, 200_000])
def test_face_flux_bitwise(V, O, dtype, nf):
    _log(f"face_flux dtype={np.dtype(dtype).name} nf={nf}")
    WL, WR, ax = random_faces(nf, seed=nf, dtype=dtype)
    F = np.zeros_like(WL)
    V.compute_face_fluxes(V.HIPFlux(), V.IdealHydro(1.4), F, WL, WR, ax)
    Fo = O.face_flux(WL, WR, ax, 1.4)
    assert np.array_equal(F, Fo), np.abs(F - Fo).max()


def test_face_flux_empty_batch_is_noop(V):
    F = np.zeros((0, 5))
    out = V.compute_face_fluxes(V.HIPFlux(), V.IdealHydro(1.4), F, np.zeros((0, 5)), np.zeros((0, 5)),
                                np.zeros(0, dtype=np.int32))
    assert out is F


def test_face_flux_device_pointer_form(V, O):
    WL, WR, ax = random_faces(10_000, seed=5)
    bWL = V.DeviceBuffer(WL.nbytes).upload(WL)
    bWR = V.DeviceBuffer(WR.nbytes).upload(WR)
    bax = V.DeviceBuffer(ax.nbytes).upload(ax)
    bF = V.DeviceBuffer(WL.nbytes)
    L = V.load_library()
    V.check(L.vh_face_flux_dev_f64(bWL.ptr, bWR.ptr, bax.ptr, len(ax), 1.4, bF.ptr, None))
    F = bF.download(np.empty_like(WL))          # hipMemcpy orders after the null-stream kernel
    assert np.array_equal(F, O.face_flux(WL, WR, ax, 1.4))
    for b in (bWL, bWR, bax, bF):
        b.free()


# --------------------------------------------------------------------------
# tier 2: fused stage kernels on a uniform grid

CASES = [
    ((400,), [(1, 1)]),                       # Sod-sized, partial tile
    ((64,), [(0, 0)]),
    ((37,), [(2, 1)]),
    ((48, 40), [(0, 0), (0, 0)]),
    ((35, 19), [(2, 1), (1, 2)]),             # partial tiles in x and y
    ((33, 9), [(0, 0), (2, 2)]),
    ((32, 16, 12), [(0, 0)] * 3),
    ((33, 17, 11), [(1, 2), (0, 0), (2, 1)]),  # everything ragged
    ((16, 16, 40), [(2, 2), (1, 1), (0, 0)]),  # several z chunks
    ((8, 8, 3), [(0, 0), (1, 1), (2, 2)]),
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_step_bitwise_vs_oracle(V, O, case, dtype):
    dims, bcs = CASES[case]
    _log(f"step dims={dims} bcs={bcs} dtype={np.dtype(dtype).name}")
    U = random_state(tuple(reversed(dims)), seed=case, dtype=dtype)
    sim = make_sim(V, dims, bcs, 1.4, dtype, U)
    g = oracle_grid(O, dims, bcs, 1.4)
    Uo = U.copy()
    f64 = dtype == np.float32
    for k in range(3):
        dt = O.compute_dt(g, Uo, 0.4, floor64=f64)
        assert V.compute_dt(sim) == dt
        V.step(sim, dt)
        O.step(g, Uo, dt, floor64=f64)
    Ug = sim.state()
    assert np.array_equal(Ug, Uo), f"max diff {np.abs(Ug.astype(float) - Uo).max()}"
    sim.close()


@pytest.mark.parametrize("dims,bcs", [((48, 40, 36), [(0, 0)] * 3), ((70, 33), [(1, 1), (2, 2)])])
def test_all_launch_shapes_agree(V, O, dims, bcs):
    _log(f"launch shapes dims={dims}")
    U = smooth_state(dims, seed=9)
    g = oracle_grid(O, dims, bcs, 5.0 / 3.0)
    Uo = U.copy()
    O.step(g, Uo, 1e-3)
    sim = make_sim(V, dims, bcs, 5.0 / 3.0, np.float64)
    zchunks = (0, 1, 5, 16, 64) if len(dims) == 3 else (0,)
    for v in range(sim.tile_count()):
        for zc in zchunks:
            sim.set_launch(v, zc)
            sim.upload(U)
            V.step(sim, 1e-3)
            assert np.array_equal(sim.state(), Uo), (v, zc, sim.launch())
    sim.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_device_dt_loop_bitwise(V, O, dtype):
    dims = (40, 24, 20)
    bcs = [(0, 0), (2, 2), (1, 1)]
    _log(f"device dt loop dtype={np.dtype(dtype).name}")
    U = smooth_state(dims, seed=3, dtype=dtype)
    sim = make_sim(V, dims, bcs, 5.0 / 3.0, dtype, U)
    sim.problem.cfl = 0.3
    g = oracle_grid(O, dims, bcs, 5.0 / 3.0)
    Uo = U.copy()
    dts = O.run_steps(g, Uo, 6, 0.3, floor64=dtype == np.float32)
    V.run_steps(sim, 6)
    assert sim.t == pytest.approx(dts.sum(), rel=1e-15)
    assert np.array_equal(sim.state(), Uo)
    sim.close()


def test_sod_400_full_run_bitwise_and_exact(V, O):
    """Config: Sod, 400 cells, outflow, CFL 0.4, to t = 0.2 — GPU ≡ oracle bit for
    bit, L1(ρ) against the exact solution under the reference's 0.01 bar."""
    _log("sod 400 full run")
    prob = V.sod_problem(400)
    sim = V.Simulation(V.UniformMesh(prob.dims, prob.domain), prob)
    U0 = sim.state().copy()
    t0 = V.conserved_totals(sim)
    V.evolve(sim)
    g = O.make_grid((400,), ((0.0, 1.0),), "outflow", 1.4)
    Uo = U0.copy()
    t, k = O.evolve(g, Uo, 0.2, 0.4)
    assert sim.t == t and sim.step_count == k
    Ug = sim.state()
    assert np.array_equal(Ug, Uo)
    t1 = V.conserved_totals(sim)
    assert t1["mass"] == pytest.approx(t0["mass"], rel=1e-12)
    assert t1["energy"] == pytest.approx(t0["energy"], rel=1e-12)
    x = (np.arange(1, 401) - 0.5) / 400
    ex = np.array([O.exact_riemann_sample((1, 0, 1), (0.125, 0, 0.1), 1.4, (xi - 0.5) / t)[0] for xi in x])
    assert np.abs(Ug[0] - ex).mean() < 0.01


def _field_rel(a, b):
    """Per-field max |a − b| / max |b| (the 1e-12 relative bar of the fast profile)."""
    a = a.reshape(5, -1).astype(np.float64)
    b = b.reshape(5, -1).astype(np.float64)
    return np.abs(a - b).max(axis=1) / np.maximum(np.abs(b).max(axis=1), 1e-300)


FAST_TOL = 1e-12


@pytest.mark.parametrize("case", [1, 3, 5, 6, 7, 8])
def test_fast_profile_within_1e12_of_oracle(V, O, case):
    dims, bcs = CASES[case]
    _log(f"fast profile dims={dims} bcs={bcs}")
    U = random_state(tuple(reversed(dims)), seed=case, dtype=np.float64, vmax=0.3)
    sim = make_sim(V, dims, bcs, 1.4, np.float64, U)
    sim.set_mode("fast")
    assert sim.mode == "fast"
    g = oracle_grid(O, dims, bcs, 1.4)
    Uo = U.copy()
    for k in range(3):
        dt = O.compute_dt(g, Uo, 0.4)
        V.step(sim, dt)
        O.step(g, Uo, dt)
    rel = _field_rel(sim.state(), Uo)
    _log(f"  per-field rel diff {rel}")
    assert (rel <= FAST_TOL).all(), rel
    sim.close()


def test_fast_profile_sod_400(V, O):
    """Fast profile through the whole Sod run (shocks, limiter switches)."""
    prob = V.sod_problem(400)
    sim = V.Simulation(V.UniformMesh(prob.dims, prob.domain), prob)
    sim.set_mode("fast")
    U0 = sim.state().copy()
    V.evolve(sim)
    g = O.make_grid((400,), ((0.0, 1.0),), "outflow", 1.4)
    t, k = O.evolve(g, U0, 0.2, 0.4)
    assert sim.step_count == k and sim.t == pytest.approx(t, rel=1e-12)
    rel = _field_rel(sim.state(), U0)
    _log(f"fast Sod-400 per-field rel diff {rel}")
    assert (rel <= 1e-10).all(), rel
    sim.close()


def test_totals_match_oracle(V, O):
    dims = (30, 20, 10)
    U = random_state((10, 20, 30), seed=2)
    sim = make_sim(V, dims, [(0, 0)] * 3, 1.4, np.float64, U)
    got = V.conserved_totals(sim)
    want = O.totals(oracle_grid(O, dims, [(0, 0)] * 3, 1.4), U)
    np.testing.assert_allclose([got[k] for k in ("mass", "momentum_x", "momentum_y", "momentum_z", "energy")],
                               want, rtol=1e-13, atol=1e-15)
    sim.close()


def test_error_paths(V):
    prob = V.Problem("bad", (1, 8), ((0, 1), (0, 1)), 1.4, "periodic", lambda x, y, z: (1, 0, 0, 0, 1), 1.0)
    with pytest.raises(V.VespaHIPError, match="periodic axis needs at least 2"):
        V.Simulation(V.UniformMesh(prob.dims, prob.domain), prob)
    with pytest.raises(ValueError):
        V.compute_face_fluxes(V.HIPFlux(), V.IdealHydro(1.4), np.zeros((2, 5)), np.zeros((2, 5)),
                              np.zeros((2, 5)), np.array([1, 4], dtype=np.int32))


# --------------------------------------------------------------------------
# full-size workload (3D 256³ periodic turbulence-like state)

@pytest.fixture(scope="module")
def turb256(V):
    import vespa_jl_amd.problems as P
    return P.turbulence_state(256)


def test_256cubed_one_step_bitwise_vs_oracle(V, O, turb256):
    _log("256^3 one step vs OpenMP oracle")
    dims = (256, 256, 256)
    g = O.make_grid(dims, [(0.0, 1.0)] * 3, "periodic", 5.0 / 3.0)
    sim = make_sim(V, dims, [(0, 0)] * 3, 5.0 / 3.0, np.float64, turb256)
    sim.problem.cfl = 0.3
    dt = O.compute_dt_omp(g, turb256, 0.3, threads=16)
    assert V.compute_dt(sim) == dt
    Uo = turb256.copy()
    O.step_omp(g, Uo, dt, threads=16)
    V.step(sim, dt)
    Ug = sim.state()
    same = np.array_equal(Ug, Uo)
    if not same:
        rel = np.abs(Ug - Uo) / np.maximum(np.abs(Uo), 1e-300)
        pytest.fail(f"not bit-identical: max rel diff {rel.max():.3e}")
    sim.close()


def test_256cubed_fast_profile(V, O, turb256):
    """Fast profile at the full size: one step against the oracle, and 100
    device-dt steps against the exact profile (which is bit-identical to the
    oracle step by step), both within 1e-12 relative per field."""
    _log("256^3 fast profile: 1 step vs oracle, 100 steps vs exact")
    dims = (256, 256, 256)
    g = O.make_grid(dims, [(0.0, 1.0)] * 3, "periodic", 5.0 / 3.0)
    dt = O.compute_dt_omp(g, turb256, 0.3, threads=16)
    Uo = turb256.copy()
    O.step_omp(g, Uo, dt, threads=16)
    fast = make_sim(V, dims, [(0, 0)] * 3, 5.0 / 3.0, np.float64, turb256)
    fast.problem.cfl = 0.3
    fast.set_mode("fast")
    V.step(fast, dt)
    rel1 = _field_rel(fast.state(), Uo)
    _log(f"  1 step: {rel1}")
    assert (rel1 <= FAST_TOL).all()
    del Uo
    exact = make_sim(V, dims, [(0, 0)] * 3, 5.0 / 3.0, np.float64, turb256)
    exact.problem.cfl = 0.3
    fast.upload(turb256)
    fast.t = 0.0
    V.run_steps(fast, 100)
    V.run_steps(exact, 100)
    assert fast.t == pytest.approx(exact.t, rel=1e-12)
    rel = _field_rel(fast.state(), exact.state())
    _log(f"  100 steps: {rel}")
    assert (rel <= FAST_TOL).all()
    fast.close()
    exact.close()


def test_256cubed_conservation_and_fp32_error(V, O, turb256):
    _log("256^3 conservation over 10 device-dt steps, fp64 and fp32")
    dims = (256, 256, 256)
    res = {}
    for dtype in (np.float64, np.float32):
        sim = make_sim(V, dims, [(0, 0)] * 3, 5.0 / 3.0, dtype, turb256.astype(dtype))
        sim.problem.cfl = 0.3
        t0 = V.conserved_totals(sim)
        V.run_steps(sim, 10)
        t1 = V.conserved_totals(sim)
        tol = 1e-13 if dtype == np.float64 else 2e-6
        assert t1["mass"] == pytest.approx(t0["mass"], rel=tol)
        assert t1["energy"] == pytest.approx(t0["energy"], rel=tol)
        for k in ("momentum_x", "momentum_y", "momentum_z"):
            assert abs(t1[k] - t0[k]) < (1e-13 if dtype == np.float64 else 1e-6)
        res[dtype] = (sim.state().astype(np.float64), sim.t)
        sim.close()
    U64, t64 = res[np.float64]
    U32, t32 = res[np.float32]
    assert t32 == pytest.approx(t64, rel=1e-5)
    l1 = np.abs(U32 - U64).mean(axis=(1, 2, 3)) / np.abs(U64).mean(axis=(1, 2, 3))
    _log(f"FP32 relative L1 vs FP64 after 10 steps per field: {l1}")
    assert (l1 < 1e-5).all()
