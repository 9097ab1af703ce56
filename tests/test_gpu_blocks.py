"""GPU parity of the batched block path (one persistent launch per RK stage over
every block, link faces between blocks, flux registers) against the block-mesh
oracle, through the C ABI.

Exact profile: bit-identical to the oracle (the reference's face walk and
evaluation order, src/driver.jl:209-359 over HGBackend's enumeration,
lib/HGBackend/src/HGBackend.jl:202-276). Fast profile: within 1e-12 relative.
"""
import numpy as np
import pytest

from conftest import load_vespa

V = load_vespa()
from vespa_jl_amd import blockmesh as BM  # noqa: E402

pytestmark = pytest.mark.gpu


def two_level_2d(bc=2):
    return BM.BlockMesh(2, 1, (1.0 / 16, 1.0 / 16), 1.4, [0, 1, 1], [[8, 8, 1], [8, 8, 1], [8, 4, 1]],
                        [[0, 0, 0], [16, 4, 0], [16, 0, 0]], [[bc] * 4 + [1, 1]] * 3)


def two_level_3d(bc=2):
    return BM.BlockMesh(3, 1, (1.0 / 8,) * 3, 1.4, [0, 1, 0], [[4, 4, 4], [8, 8, 8], [4, 4, 4]],
                        [[0, 0, 0], [8, 0, 0], [16, 0, 0]], [[bc] * 6] * 3)


def fine_below_3d():
    """Block 1 (coarse) has a finer block 0 on its lo-y side and a same-level block
    2 on its lo-x side: foreign groups before and after its own contributions."""
    return BM.BlockMesh(3, 1, (1.0 / 8,) * 3, 1.4, [1, 0, 0], [[8, 4, 8], [4, 4, 4], [4, 4, 4]],
                        [[8, 0, 0], [8, 4, 0], [0, 4, 0]], [[1, 2, 2, 1, 1, 2]] * 3)


def sod_blocks():
    return BM.BlockMesh(1, 1, (1.0 / 64,), 1.4, [0, 1, 0], [[16, 1, 1], [64, 1, 1], [16, 1, 1]],
                        [[0, 0, 0], [32, 0, 0], [96, 0, 0]], [[1] * 6] * 3)


def random_mesh(n=14, seed=3, rank=3, edges=(8, 16, 24)):
    return BM.random_block_pairs(n, seed=seed, rank=rank, edges=edges)


MESHES = {
    "2d_two_level": two_level_2d,
    "3d_two_level": two_level_3d,
    "3d_fine_below": fine_below_3d,
    "1d_sod": sod_blocks,
    "3d_random": random_mesh,
    "2d_random": lambda: random_mesh(16, 5, 2, (8, 16, 24, 32)),
}


def _state(m, dtype, seed=1):
    return BM.smooth_block_state(m, seed=seed).astype(dtype)


def _sim(m, dtype, mode="exact", **launch):
    s = BM.BlockSimulation(m, dtype, cfl=0.3)
    s.set_mode(mode)
    if launch:
        s.set_launch(**launch)
    return s


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("name", sorted(MESHES))
def test_step_bitwise(O, name, dtype):
    m = MESHES[name]()
    print(f"blocks step {name} {np.dtype(dtype).name}", flush=True)
    f64 = dtype == np.float64
    bo = O.BlockOracle(m)
    U = _state(m, dtype)
    s = _sim(m, dtype)
    s.upload(U)
    for k in range(3):
        dt = bo.compute_dt(U, 0.3, floor64=not f64)
        assert s.compute_dt() == dt
        bo.step(U, dt, floor64=not f64)
        s.step(dt)
        got = s.state()
        np.testing.assert_array_equal(got, U, err_msg=f"step {k}")
    s.close()


@pytest.mark.parametrize("code,dims", [(0, (12, 10, 8)), (1, (9, 7, 5)), (2, (16, 6)), (0, (40,))])
def test_single_block_equals_uniform_grid(O, code, dims):
    """One block is RefMesh: the block path equals the oracle and the uniform-grid
    path bit for bit."""
    rank = len(dims)
    m = BM.uniform_as_blocks(dims, [(0.0, 1.0)] * rank, [(code, code)] * rank, 1.4)
    print(f"single block {dims} bc {code}", flush=True)
    shape = tuple(reversed(dims))
    from helpers import random_state
    U = random_state(shape, seed=sum(dims)).reshape(5, -1).copy()
    bo = O.BlockOracle(m)
    s = _sim(m, np.float64)
    s.upload(U)
    names = {0: "periodic", 1: "outflow", 2: "reflecting"}
    mesh = V.UniformMesh(tuple(dims), tuple((0.0, 1.0) for _ in dims))
    prob = V.Problem("one block", tuple(dims), tuple((0.0, 1.0) for _ in dims), 1.4, names[code],
                     lambda x, y, z: (1.0, 0.0, 0.0, 0.0, 1.0), 1.0, 0.3)
    g = V.Simulation(mesh, prob)
    g.upload(U.reshape((5,) + shape))
    for _ in range(2):
        dt = bo.compute_dt(U, 0.3)
        bo.step(U, dt)
        s.step(dt)
        V.step(g, dt)
    np.testing.assert_array_equal(s.state(), U)
    np.testing.assert_array_equal(g.state().reshape(5, -1), U)
    s.close()
    g.close()


@pytest.mark.parametrize("level", [0, 1])
@pytest.mark.parametrize("name", ["2d_two_level", "3d_two_level", "3d_random"])
def test_level_restricted_step_bitwise(O, name, level):
    m = MESHES[name]()
    print(f"level step {name} L{level}", flush=True)
    bo = O.BlockOracle(m)
    U = _state(m, np.float64, seed=2)
    s = _sim(m, np.float64)
    s.upload(U)
    dt = bo.compute_dt(U, 0.3, level=level)
    assert s.compute_dt(level=level) == dt
    bo.step(U, dt, level=level)
    s.step(dt, level=level)
    np.testing.assert_array_equal(s.state(), U)
    s.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("name", ["2d_two_level", "3d_two_level", "1d_sod", "3d_fine_below"])
def test_subcycled_evolve_bitwise_and_conservative(O, name, dtype):
    """evolve_level! recursion with flux registers (src/driver.jl:553-605,
    src/reflux.jl:51-103): bitwise with the oracle; closed boxes conserve."""
    m = MESHES[name]()
    print(f"subcycle {name} {np.dtype(dtype).name}", flush=True)
    f64 = dtype == np.float64
    bo = O.BlockOracle(m)
    U = _state(m, dtype, seed=4)
    s = _sim(m, dtype)
    s.upload(U)
    tf = 0.01
    before = s.totals()
    t_o, k_o = bo.evolve(U, tf, 0.3, subcycle=True, floor64=not f64)
    t_g, k_g = s.evolve(tf, subcycle=True)
    assert (t_g, k_g) == (t_o, k_o)
    np.testing.assert_array_equal(s.state(), U)
    if name in ("2d_two_level", "3d_two_level"):         # reflecting boxes
        after = s.totals()
        rel = 1e-13 if f64 else 1e-6
        assert after[0] == pytest.approx(before[0], rel=rel)
        assert after[4] == pytest.approx(before[4], rel=rel)
    s.close()


@pytest.mark.parametrize("name", ["3d_random", "2d_two_level", "1d_sod"])
def test_single_rate_evolve_bitwise(O, name):
    """Device dt loop (fused CFL partials, graph replays) against evolve!."""
    m = MESHES[name]()
    print(f"single-rate evolve {name}", flush=True)
    bo = O.BlockOracle(m)
    U = _state(m, np.float64, seed=5)
    s = _sim(m, np.float64)
    s.upload(U)
    t_o, k_o = bo.evolve(U, 0.02, 0.3)
    t_g, k_g = s.evolve(0.02)
    assert k_g == k_o and t_g == t_o
    np.testing.assert_array_equal(s.state(), U)
    s.close()


def test_launch_shapes_and_methods_agree(O):
    """zchunk, the workgroup cap, graphs and the per-block launch loop all give
    the same bits as the oracle."""
    m = random_mesh(12, seed=9, rank=3, edges=(8, 16, 24, 32))
    print("launch shapes", flush=True)
    bo = O.BlockOracle(m)
    U0 = _state(m, np.float64, seed=6)
    U = U0.copy()
    for _ in range(3):
        bo.step(U, bo.compute_dt(U, 0.3))
    for launch in ({"zchunk": 0}, {"zchunk": 8}, {"zchunk": 5, "max_workgroups": 7},
                   {"zchunk": 16, "use_graph": False}):
        s = _sim(m, np.float64, **launch)
        s.upload(U0)
        s.evolve(max_steps=3)
        np.testing.assert_array_equal(s.state(), U, err_msg=str(launch))
        s.close()
    for method in ("persistent", "per_block", "graph"):
        s = _sim(m, np.float64)
        s.upload(U0)
        r = s.time_steps(1, method=method, stage2=False)
        assert r["step_ms"] > 0
        # two untimed steps + one timed step
        np.testing.assert_array_equal(s.state(), U, err_msg=method)
        s.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_fast_profile_within_1e12(O, dtype):
    m = random_mesh(12, seed=10, rank=3, edges=(8, 16))
    print(f"fast profile {np.dtype(dtype).name}", flush=True)
    U0 = _state(m, dtype, seed=7)
    out = {}
    for mode in ("exact", "fast"):
        s = _sim(m, dtype, mode)
        s.upload(U0)
        s.evolve(max_steps=5)
        out[mode] = s.state().astype(np.float64)
        s.close()
    scale = np.abs(out["exact"]).max(axis=1, keepdims=True)
    tol = 1e-12 if dtype == np.float64 else 2e-6
    assert (np.abs(out["fast"] - out["exact"]) / scale).max() <= tol


def test_totals_and_conservation(O):
    m = random_mesh(12, seed=8, rank=3, edges=(8, 16))
    m.bc[:] = 2
    print("totals", flush=True)
    bo = O.BlockOracle(m)
    U = _state(m, np.float64, seed=8)
    s = _sim(m, np.float64)
    s.upload(U)
    t0 = s.totals()
    np.testing.assert_allclose(t0, bo.totals(U), rtol=1e-13, atol=1e-15)
    s.evolve(max_steps=10)
    t1 = s.totals()
    assert t1[0] == pytest.approx(t0[0], rel=1e-13)
    assert t1[4] == pytest.approx(t0[4], rel=1e-13)
    s.close()


def test_bad_mesh_raises_before_touching_the_device():
    m = BM.BlockMesh(2, 2, (0.1, 0.1), 1.4, [0, 2], [[4, 4, 1]] * 2, [[0, 0, 0], [16, 0, 0]], [[1] * 6] * 2)
    with pytest.raises(V.VespaHIPError, match="more than one level"):
        BM.BlockSimulation(m)


def test_c3_full_size_sampled_units_bitwise(O):
    """The 3,000-block workload at full size (37 M cells): the device CFL dt of
    the whole table equals the oracle's, and after two steps every sampled
    independent unit (blocks connected by faces) equals the oracle run on that
    unit alone with the same dt, bit for bit."""
    m = BM.random_block_pairs(3000, seed=2024, rank=3, edges=(8, 16, 24, 32))
    print(f"C3 full size: {m.ncell} cells", flush=True)
    U0 = BM.smooth_block_state(m, seed=1)
    s = _sim(m, np.float64)
    s.upload(U0)
    dt0 = s.compute_dt()
    assert dt0 == O.BlockOracle(m).compute_dt(U0, 0.3)
    dts = []
    for _ in range(2):
        dt = s.compute_dt()
        s.step(dt)
        dts.append(dt)
    G = s.state()
    s.close()
    comps = BM.components(m)
    rng = np.random.default_rng(0)
    pairs = [g for g in comps if len(g) == 2 and m.level[g].max() == 1]
    conf = [g for g in comps if len(g) == 2 and m.level[g].max() == 0]
    single = [g for g in comps if len(g) == 1]
    picks = [pairs[i] for i in rng.choice(len(pairs), 12, replace=False)]
    picks += [conf[i] for i in rng.choice(len(conf), 6, replace=False)]
    picks += [single[i] for i in rng.choice(len(single), 4, replace=False)]
    for g in picks:
        sub, cells = BM.submesh(m, g)
        bo = O.BlockOracle(sub)
        Ug = np.ascontiguousarray(U0[:, cells])
        for dt in dts:
            bo.step(Ug, dt)
        np.testing.assert_array_equal(G[:, cells], Ug, err_msg=f"unit {g}")
