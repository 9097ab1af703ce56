"""GPU parity of the device regrid against the oracle, through the C ABI:
density_gradient_indicator per leaf and per block (src/driver.jl:400-414), the
injection remap onto a refined mesh (lib/RefMesh/src/RefMesh.jl:284-302), and a
whole dynamic-AMR Sedov run whose regrids happen on the device, replayed on the
CPU by the oracle (indicator, tagging, refinement, remap, evolve!). Bit-exact:
the indicator and the remap are integer/gather work plus one subtraction and
one division per neighbour; the blocks stage kernel in the exact profile is
bit-identical to the oracle."""
import numpy as np
import pytest

from conftest import load_vespa

V = load_vespa()
from vespa_jl_amd import amr, blockmesh as BM  # noqa: E402
from vespa_jl_amd._lib import VespaHIPError  # noqa: E402

pytestmark = pytest.mark.gpu


def _bumped_quadblocks(nreg=3):
    """A QuadBlocks mesh refined around a density bump (levels 0..3)."""
    qb = amr.QuadBlocks(6, 8, 3)
    for _ in range(nreg):
        x, y = qb.cell_centers()
        rho = 1.0 + 0.5 * np.exp(-(x * x + y * y) / 0.01)
        g = np.zeros(len(qb.blocks))
        for q in range(len(qb.blocks)):
            r = rho[q * 64:(q + 1) * 64]
            g[q] = (r.max() - r.min()) / r.max()
        qb.refine(qb.tags_from_indicator(g, 0.02, symmetric=True))
    return qb


def _split_all(mesh, pick):
    """Refine the blocks b with pick(b) into 2^rank children of the same cell
    counts (dims must be even); returns the new mesh and the source table."""
    lv, dm, og, bc, src = [], [], [], [], []
    for b in range(mesh.nblocks):
        if not pick(b):
            lv.append(int(mesh.level[b])); dm.append(mesh.dims[b].tolist()); og.append(mesh.origin[b].tolist())
            bc.append(mesh.bc[b].tolist()); src.append(b)
            continue
        s_child = 1 << (mesh.maxlevel - int(mesh.level[b]) - 1)
        n = mesh.dims[b]
        for cz in range(2 if mesh.rank > 2 else 1):
            for cy in range(2 if mesh.rank > 1 else 1):
                for cx in range(2):
                    c = (cx, cy, cz)
                    lv.append(int(mesh.level[b]) + 1)
                    dm.append(n.tolist())
                    og.append([int(mesh.origin[b][d]) + c[d] * int(n[d]) * s_child if d < mesh.rank else 0
                               for d in range(3)])
                    bc.append(mesh.bc[b].tolist())
                    src.append(b)
    m = BM.BlockMesh(mesh.rank, mesh.maxlevel, mesh.h0[:mesh.rank], mesh.gamma, lv, dm, og, bc)
    return m, np.array(src, dtype=np.int32)


INDICATOR_MESHES = {
    "quadblocks": lambda: _bumped_quadblocks().mesh(),
    "2d_random": lambda: BM.random_block_pairs(16, seed=5, rank=2, edges=(8, 16, 24, 32)),
    "3d_random": lambda: BM.random_block_pairs(14, seed=3, rank=3, edges=(8, 16, 24)),
    "periodic_block": lambda: BM.uniform_as_blocks((24, 20), [(0.0, 1.0)] * 2, [(0, 0), (2, 2)], 1.4),
}


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("name", sorted(INDICATOR_MESHES))
def test_indicator_bitwise(O, name, dtype):
    print(f"indicator {name} {np.dtype(dtype).name}", flush=True)
    m = INDICATOR_MESHES[name]()
    U = BM.smooth_block_state(m, seed=7, dtype=dtype)
    bmax_o, cell_o = O.BlockOracle(m).indicator(U)
    s = BM.BlockSimulation(m, dtype)
    try:
        s.upload(U)
        bmax, cell = s.indicator(cells=True)
        assert np.array_equal(cell, cell_o)
        assert np.array_equal(bmax, bmax_o)
        assert np.array_equal(s.indicator(), bmax_o)
    finally:
        s.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_remap_quadblocks_bitwise(O, dtype):
    """Three regrid rounds of the quadtree driver: device remap == oracle remap."""
    print(f"remap quadblocks {np.dtype(dtype).name}", flush=True)
    qb = amr.QuadBlocks(6, 8, 3)
    U = amr.sedov_quadblocks_state(qb).astype(dtype)
    x, y = qb.cell_centers()
    U[0] += (0.5 * np.exp(-(x * x + y * y) / 0.01)).astype(dtype)
    s = BM.BlockSimulation(qb.mesh(), dtype)
    s.upload(U)
    try:
        for r in range(3):
            old, old_mesh = list(qb.blocks), qb.mesh()
            qb.refine(qb.tags_from_indicator(s.indicator(), 0.02, symmetric=True))
            src = qb.source_blocks(old)
            U = O.blocks_remap(qb.mesh(), old_mesh, src, U)
            new = BM.BlockSimulation(qb.mesh(), dtype)
            new.remap_from(s, src)
            s.close()
            s = new
            assert np.array_equal(s.state(), U), f"round {r}"
        assert max(b[0] for b in qb.blocks) == 3
    finally:
        s.close()


@pytest.mark.parametrize("rank", [1, 2, 3])
def test_remap_split_blocks_bitwise(O, rank):
    """Every other block of a random mesh split into 2^rank children (rank 1:
    the 1D Sod blocks)."""
    print(f"remap split rank {rank}", flush=True)
    if rank == 1:
        m = BM.BlockMesh(1, 2, (1.0 / 64,), 1.4, [0, 1, 0], [[16, 1, 1], [64, 1, 1], [16, 1, 1]],
                         [[0, 0, 0], [64, 0, 0], [192, 0, 0]], [[1] * 6] * 3)
    else:
        m = BM.random_block_pairs(8, seed=11, rank=rank, edges=(8, 16))
        m = BM.BlockMesh(rank, 2, m.h0[:rank], m.gamma, m.level, m.dims, m.origin * 2, m.bc)
    m2, src = _split_all(m, lambda b: m.level[b] == 0 and b % 2 == 0)
    U = BM.smooth_block_state(m, seed=2)
    a = BM.BlockSimulation(m, np.float64)
    b = BM.BlockSimulation(m2, np.float64)
    try:
        a.upload(U)
        b.remap_from(a, src)
        assert np.array_equal(b.state(), O.blocks_remap(m2, m, src, U))
    finally:
        a.close()
        b.close()


def test_remap_rejects_bad_tables():
    qb = amr.QuadBlocks(2, 8, 1)
    a = BM.BlockSimulation(qb.mesh(), np.float64)
    old = list(qb.blocks)
    qb.refine({qb.blocks[0]})
    b = BM.BlockSimulation(qb.mesh(), np.float64)
    c = BM.BlockSimulation(qb.mesh(), np.float32)
    try:
        good = qb.source_blocks(old)
        with pytest.raises(ValueError):
            b.remap_from(a, good[:-1])
        bad = good.copy()
        bad[0] = 99
        with pytest.raises(VespaHIPError, match="out of range"):
            b.remap_from(a, bad)
        bad = good.copy()
        bad[-1] = (bad[-1] + 1) % len(old)          # a level-1 block whose "parent" does not contain it
        with pytest.raises(VespaHIPError):
            b.remap_from(a, bad)
        with pytest.raises(VespaHIPError, match="coarser"):
            a.remap_from(b, np.zeros(len(old), dtype=np.int32) + len(qb.blocks) - 1)
        with pytest.raises(VespaHIPError, match="dtype"):
            c.remap_from(a, good)
        b.remap_from(a, good)                          # still usable after the rejections
    finally:
        a.close()
        b.close()
        c.close()


def _oracle_sedov(O, nroot, B, maxlevel, tfinal, every, dtype, cfl=0.4, thr=0.1):
    """run_sedov restated on the CPU: oracle indicator, the same host policy,
    oracle remap, oracle evolve! with the global dt."""
    qb = amr.QuadBlocks(nroot, B, maxlevel)
    U = np.ascontiguousarray(amr.sedov_quadblocks_state(qb).astype(dtype))
    t, steps, nreg = 0.0, 0, 0
    while t < tfinal * (1 - 1e-12):
        mesh = qb.mesh()
        bmax, _ = O.BlockOracle(mesh).indicator(U)
        tags = qb.tags_from_indicator(bmax, thr, symmetric=True)
        if tags:
            old = list(qb.blocks)
            qb.refine(tags)
            U = np.ascontiguousarray(O.blocks_remap(qb.mesh(), mesh, qb.source_blocks(old), U))
            nreg += 1
        t, k = O.BlockOracle(qb.mesh()).evolve(U, tfinal, cfl, max_steps=every, floor64=dtype == np.float32, t0=t)
        steps += k
    return qb, U, t, steps, nreg


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sedov_run_with_device_regrid_bitwise(O, dtype):
    print(f"sedov device regrid {np.dtype(dtype).name}", flush=True)
    kw = dict(nroot=6, B=8, maxlevel=3, tfinal=0.012, every=4)
    qb, U, t, steps, nreg = _oracle_sedov(O, dtype=dtype, **kw)
    r = amr.run_sedov(dtype=dtype, mode="exact", **kw)
    assert nreg >= 4 and r["regrids"] == nreg
    assert r["blocks"] == qb.blocks
    assert r["t"] == t and r["steps"] == steps
    assert np.array_equal(r["U"], U)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_regrid_in_place_equals_new_context(O, dtype):
    """vh_blocks_regrid (regrid! in place) against a fresh context of the new
    mesh filled by vh_blocks_remap and against the oracle's remap: the same
    state bit for bit after each of several rounds (the first outgrows the
    buffers, later small rounds reuse them), and the same steps afterwards, in
    the profile and launch shape the context had before."""
    print(f"regrid in place {np.dtype(dtype).name}", flush=True)
    qb = amr.QuadBlocks(6, 8, 3)
    U = amr.sedov_quadblocks_state(qb).astype(dtype)
    x, y = qb.cell_centers()
    U[0] += (0.5 * np.exp(-(x * x + y * y) / 0.01)).astype(dtype)
    s = BM.BlockSimulation(qb.mesh(), dtype)
    s.set_mode("exact")
    s.set_launch(zchunk=0, max_workgroups=0, use_graph=True)
    s.upload(U)
    try:
        s.evolve(max_steps=2)
        for r in range(4):
            Ub, old, old_mesh = s.state(), list(qb.blocks), qb.mesh()
            tags = qb.tags_from_indicator(s.indicator(), 0.02 if r == 0 else 0.3, symmetric=True)
            if not tags:
                tags = {next(b for b in qb.blocks if b[0] < qb.maxlevel)}
            qb.refine(tags)
            src = qb.source_blocks(old)
            ref = BM.BlockSimulation(qb.mesh(), dtype)
            ref.set_mode("exact")
            try:
                ref.remap_from(s, src)
                s.regrid(qb.mesh(), src)
                assert s.ncell == ref.ncell == qb.mesh().ncell
                assert np.array_equal(s.state(), O.blocks_remap(qb.mesh(), old_mesh, src, Ub)), f"round {r}"
                assert np.array_equal(s.state(), ref.state()), f"round {r}"
                s.evolve(max_steps=2)
                ref.evolve(max_steps=2)
                assert s.t == ref.t
                assert np.array_equal(s.state(), ref.state()), f"round {r} steps"
            finally:
                ref.close()
    finally:
        s.close()


def test_regrid_in_place_with_gravity_rebuilds_the_operator(O):
    """With self-gravity on, the regrid rebuilds the composite operator for the
    new mesh (A·x equals a fresh context's and the oracle's) and carries φ over
    by the same injection as the state."""
    print("regrid in place with gravity", flush=True)
    qb = amr.QuadBlocks(4, 8, 2, bc=BM.REF)
    x, y = qb.cell_centers()
    U = np.zeros((5, len(x)))
    U[0] = 1 + 2.0 * np.exp(-(x * x + y * y) / 0.02)
    U[4] = 1.0 / 0.4
    s = BM.BlockSimulation(qb.mesh(), np.float64)
    try:
        s.upload(U)
        s.enable_gravity(G=1.0, maxiter=2000, tol=1e-10)
        s.solve_poisson()
        phi0 = s.phi
        old, old_mesh = list(qb.blocks), qb.mesh()
        qb.refine({(0, 1, 1), (0, 2, 1), (0, 1, 2), (0, 2, 2)})
        src = qb.source_blocks(old)
        s.regrid(qb.mesh(), src)
        m = qb.mesh()
        phi_expect = O.blocks_remap(m, old_mesh, src, phi0[None, :])[0]
        assert np.array_equal(s.phi, phi_expect)
        xr = np.random.default_rng(3).normal(size=m.ncell)
        assert np.array_equal(s.apply_laplacian(xr), O.BlockOracle(m).poisson_apply(xr))
        it, rr = s.solve_poisson()
        assert rr <= 1e-10
        s.evolve(max_steps=2)
    finally:
        s.close()


def test_regrid_rejects_bad_tables_and_keeps_the_context():
    qb = amr.QuadBlocks(2, 8, 1)
    s = BM.BlockSimulation(qb.mesh(), np.float64)
    try:
        x, y = qb.cell_centers()
        U0 = np.stack([1 + 0.1 * x, 0.1 * y, 0 * x, 0 * x, 2 + 0 * x])
        s.upload(U0)
        old = list(qb.blocks)
        mesh0 = qb.mesh()
        qb.refine({qb.blocks[0]})
        bad = qb.source_blocks(old).copy()
        bad[0] = 99
        with pytest.raises(VespaHIPError, match="out of range"):
            s.regrid(qb.mesh(), bad)
        wrong = BM.BlockMesh(2, 2, qb.mesh().h0[:2], qb.mesh().gamma, [0], [[8, 8, 1]], [[0, 0, 0]], [[1] * 6])
        with pytest.raises(ValueError, match="maxlevel"):
            s.regrid(wrong, [0])
        assert s.ncell == mesh0.ncell and np.array_equal(s.state(), U0)   # untouched
        s.evolve(max_steps=1)
    finally:
        s.close()
