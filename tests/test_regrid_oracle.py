"""Pins of the regrid oracle (CPU): density_gradient_indicator on block meshes
(src/driver.jl:400-414) and the injection remap (prolong!,
lib/RefMesh/src/RefMesh.jl:284-302).

The indicator is pinned by known answers on one block (a density step, the
periodic wrap, the Float32 difference), and on multi-level meshes by a
brute-force neighbour search written here independently of the oracle's
topology (the point one lattice unit across the face's lo corner, located by
scanning every block). The remap is pinned by what injection fixes exactly:
the piecewise-constant field on the finest lattice is unchanged, and so is
Σ U·V."""
import numpy as np
import pytest

from conftest import load_vespa

V = load_vespa()
from vespa_jl_amd import amr, blockmesh as BM  # noqa: E402

EPS = np.finfo(np.float64).eps


def _one_block(nx, ny, code):
    return BM.uniform_as_blocks((nx, ny), [(0.0, 1.0), (0.0, 1.0)], [(code, code)] * 2, 1.4)


def _state_with_rho(rho, dtype=np.float64):
    n = rho.size
    U = np.zeros((5, n), dtype=dtype)
    U[0] = rho.ravel()
    U[4] = 1.0
    return U


def test_indicator_density_step(O):
    """ρ = 1 left of x = 5, 0.125 right: the two columns at the step see
    0.875 / 1; every other cell sees 0 (reflecting walls are not neighbours)."""
    nx, ny = 10, 6
    rho = np.where(np.arange(nx)[None, :] < 5, 1.0, 0.125) * np.ones((ny, 1))
    bo = O.BlockOracle(_one_block(nx, ny, BM.REF))
    bmax, g = bo.indicator(_state_with_rho(rho))
    g = g.reshape(ny, nx)
    expect = np.zeros((ny, nx))
    expect[:, 4] = expect[:, 5] = 0.875
    assert np.array_equal(g, expect)
    assert bmax[0] == 0.875


def test_indicator_periodic_wrap(O):
    """A density spike in column 0: with periodic x its wrap neighbour (column
    nx-1) sees it; with reflecting walls it does not."""
    nx, ny = 8, 3
    rho = np.ones((ny, nx))
    rho[:, 0] = 2.0
    gp = O.BlockOracle(_one_block(nx, ny, BM.PER)).indicator(_state_with_rho(rho))[1].reshape(ny, nx)
    gr = O.BlockOracle(_one_block(nx, ny, BM.REF)).indicator(_state_with_rho(rho))[1].reshape(ny, nx)
    expect = np.zeros((ny, nx))
    expect[:, [0, 1, nx - 1]] = 0.5
    assert np.array_equal(gp, expect)
    expect[:, nx - 1] = 0.0
    assert np.array_equal(gr, expect)


def test_indicator_float32_difference(O):
    """The difference is taken in Float32, the ratio in Float64: ρ = 1 next to
    2e-8 gives fl32(1 − 2e-8) = 1, so g = 1 exactly (the Float64 difference
    would give 1 − 2e-8)."""
    rho = np.array([[1.0, 2e-8]], dtype=np.float32)
    bo = O.BlockOracle(_one_block(2, 1, BM.REF))
    g32 = bo.indicator(_state_with_rho(rho, np.float32))[1]
    g64 = bo.indicator(_state_with_rho(rho.astype(np.float64)))[1]
    assert g32[0] == 1.0 and g32[1] == 1.0
    assert g64[0] == float(np.float64(1.0) - np.float64(np.float32(2e-8))) < 1.0


def _brute_indicator(mesh, rho):
    """Per-cell indicator by brute force: the representative neighbour across
    face (d, side) of a cell is the leaf containing the lattice point one unit
    across the face's lo corner; blocks are scanned one by one."""
    off = mesh.offsets
    out = np.zeros(mesh.ncell)
    span = [1 << (mesh.maxlevel - int(l)) for l in mesh.level]

    def locate(q):
        for c in range(mesh.nblocks):
            o, n, s = mesh.origin[c], mesh.dims[c], span[c]
            if all(o[e] <= q[e] < o[e] + n[e] * s for e in range(mesh.rank)):
                t = [(q[e] - o[e]) // s if e < mesh.rank else 0 for e in range(3)]
                return off[c] + (t[2] * n[1] + t[1]) * n[0] + t[0]
        return None

    for b in range(mesh.nblocks):
        n, s, o = mesh.dims[b], span[b], mesh.origin[b]
        for c in range(off[b + 1] - off[b]):
            ijk = (c % n[0], (c // n[0]) % n[1], c // (n[0] * n[1]))
            rc = rho[off[b] + c]
            g = 0.0
            for d in range(mesh.rank):
                for side in (0, 1):
                    j = ijk[d] + (1 if side else -1)
                    if mesh.bc[b][2 * d + side] == BM.PER and not (0 <= j < n[d]):
                        t = list(ijk)
                        t[d] = j % n[d]
                        nb = off[b] + (t[2] * n[1] + t[1]) * n[0] + t[0]
                    else:
                        q = [o[e] + ijk[e] * s for e in range(3)]
                        q[d] = q[d] + s if side else q[d] - 1
                        nb = locate(q)
                    if nb is None:
                        continue
                    rn = rho[nb]
                    g = max(g, float(abs(rn - rc)) / max(float(rc), float(rn), EPS))
            out[off[b] + c] = g
    return out


def _refined_quadblocks(steps=2, seed=0):
    qb = amr.QuadBlocks(3, 4, 2)
    rng = np.random.default_rng(seed)
    for _ in range(steps):
        bm = rng.uniform(0, 1, len(qb.blocks))
        qb.refine(qb.tags_from_indicator(bm, 0.6))
    return qb


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_indicator_matches_brute_force_on_multilevel_meshes(O, dtype):
    for mesh, seed in ((_refined_quadblocks().mesh(), 1), (BM.random_block_pairs(10, seed=5, rank=2, edges=(4, 6)), 2),
                       (BM.random_block_pairs(6, seed=7, rank=3, edges=(2, 4)), 3)):
        assert len(set(mesh.level.tolist())) > 1
        U = BM.smooth_block_state(mesh, seed, dtype)
        bmax, g = O.BlockOracle(mesh).indicator(U)
        ref = _brute_indicator(mesh, U[0])
        assert np.array_equal(g, ref)
        off = mesh.offsets
        assert np.array_equal(bmax, [g[off[b]:off[b + 1]].max() for b in range(mesh.nblocks)])


def test_remap_preserves_the_fine_raster_and_totals(O):
    """Injection leaves the field on the finest lattice unchanged, hence Σ U·V."""
    qb = amr.QuadBlocks(4, 8, 3)
    rng = np.random.default_rng(4)
    U = rng.uniform(0.5, 2.0, (5, len(qb.blocks) * 64))
    for _ in range(3):
        old_mesh, old_blocks = qb.mesh(), list(qb.blocks)
        before = qb.raster(U)
        tot = (U * qb.volumes()).sum(axis=1)
        qb.refine(qb.tags_from_indicator(rng.uniform(0, 1, len(old_blocks)), 0.7))
        U = O.blocks_remap(qb.mesh(), old_mesh, qb.source_blocks(old_blocks), U)
        assert np.array_equal(qb.raster(U), before)
        np.testing.assert_allclose((U * qb.volumes()).sum(axis=1), tot, rtol=1e-14)
    assert max(b[0] for b in qb.blocks) == 3


def test_remap_identity_and_rejects_coarsening(O):
    qb = _refined_quadblocks()
    m = qb.mesh()
    U = np.arange(5 * m.ncell, dtype=np.float64).reshape(5, -1)
    assert np.array_equal(O.blocks_remap(m, m, np.arange(m.nblocks), U), U)
    coarse = amr.QuadBlocks(3, 4, 2)
    with pytest.raises(ValueError):
        O.blocks_remap(coarse.mesh(), m, np.zeros(coarse.mesh().nblocks, dtype=int) + m.nblocks - 1,
                       U)
