"""Pins of the block-mesh oracle (hanging-node faces, flux registers, subcycling).

Anchors: a single block is RefMesh (bit-identical to the uniform oracle); the
reference's AMR thresholds (conservation rtol 1e-9 with the batched flux,
test/test_amr_ka_flux.jl:38-68; reflux restores subcycled conservation,
test/test_reflux.jl:14-55); round-off conservation in closed boxes; free-stream
preservation across 2:1 faces; Sod through a refined region against the exact
solution.
"""
import numpy as np
import pytest

from conftest import load_vespa
from helpers import random_state

V = load_vespa()
from vespa_jl_amd import blockmesh as BM  # noqa: E402

NAMES = {0: "periodic", 1: "outflow", 2: "reflecting"}


@pytest.mark.parametrize("dims,code", [((12,), 1), ((9,), 2), ((10, 8), 2), ((8, 6), 0), ((6, 5, 4), 0),
                                       ((5, 4, 6), 1)])
def test_single_block_is_refmesh_bitwise(O, dims, code):
    m = BM.uniform_as_blocks(dims, [(0.0, 1.0)] * len(dims), [(code, code)] * len(dims), 1.4)
    bo = O.BlockOracle(m)
    assert bo.validate() == 0
    shape = tuple(reversed(dims))
    U = random_state(shape, seed=sum(dims)).reshape(5, -1).copy()
    g = O.make_grid(dims, [(0.0, 1.0)] * len(dims), NAMES[code], 1.4)
    Ug = U.reshape((5,) + shape).copy()
    assert np.array_equal(bo.accumulate_flux(U), O.accumulate_flux(g, Ug).reshape(5, -1))
    dt = O.compute_dt(g, Ug, 0.4)
    assert bo.compute_dt(U, 0.4) == dt
    bo.step(U, dt)
    O.step(g, Ug, dt)
    assert np.array_equal(U, Ug.reshape(5, -1))


def test_1d_same_level_blocks_equal_uniform_bitwise(O):
    """In 1D the leaf order of three abutting blocks is the uniform order."""
    h0 = 1.0 / 48
    m = BM.BlockMesh(1, 0, (h0,), 1.4, [0, 0, 0], [[16, 1, 1]] * 3, [[0, 0, 0], [16, 0, 0], [32, 0, 0]],
                     [[1] * 6] * 3)
    bo = O.BlockOracle(m)
    U = random_state((48,), seed=3)
    Ub = U.reshape(5, -1).copy()
    g = O.make_grid((48,), ((0.0, 1.0),), "outflow", 1.4)
    for _ in range(4):
        dt = O.compute_dt(g, U, 0.4)
        assert bo.compute_dt(Ub, 0.4) == dt
        O.step(g, U, dt)
        bo.step(Ub, dt)
    assert np.array_equal(Ub, U.reshape(5, -1))


def test_3d_same_level_split_matches_uniform_to_roundoff(O):
    h0 = 1.0 / 16
    # two 8x16x16 halves of a 16^3 box along x (reflecting box)
    m = BM.BlockMesh(3, 0, (h0,) * 3, 1.4, [0, 0], [[8, 16, 16]] * 2, [[0, 0, 0], [8, 0, 0]], [[2] * 6] * 2)
    bo = O.BlockOracle(m)
    U = random_state((16, 16, 16), seed=4)
    g = O.make_grid((16, 16, 16), [(0.0, 1.0)] * 3, "reflecting", 1.4)
    Ub = np.concatenate([U[:, :, :, :8].reshape(5, -1), U[:, :, :, 8:].reshape(5, -1)], axis=1)
    dt = O.compute_dt(g, U, 0.4)
    O.step(g, U, dt)
    bo.step(Ub, dt)
    back = np.concatenate([Ub[:, :2048].reshape(5, 16, 16, 8), Ub[:, 2048:].reshape(5, 16, 16, 8)], axis=3)
    np.testing.assert_allclose(back, U, rtol=1e-13, atol=1e-14)


def two_level_2d(bc=2):
    """Coarse 8x8 block (level 0) with a fine 8x8 block (level 1) on part of its
    hi-x face, plus a coarse 4x8 block beside the fine one."""
    return BM.BlockMesh(2, 1, (1.0 / 16, 1.0 / 16), 1.4, [0, 1, 1],
                        [[8, 8, 1], [8, 8, 1], [8, 4, 1]],
                        [[0, 0, 0], [16, 4, 0], [16, 0, 0]], [[bc] * 4 + [1, 1]] * 3)


def two_level_3d(bc=2):
    return BM.BlockMesh(3, 1, (1.0 / 8,) * 3, 1.4, [0, 1, 0],
                        [[4, 4, 4], [8, 8, 8], [4, 4, 4]],
                        [[0, 0, 0], [8, 0, 0], [16, 0, 0]], [[bc] * 6] * 3)


@pytest.mark.parametrize("maker", [two_level_2d, two_level_3d])
def test_closed_two_level_mesh_conserves_to_roundoff(O, maker):
    m = maker()
    m.validate()
    bo = O.BlockOracle(m)
    assert bo.validate() == 0
    U = BM.smooth_block_state(m, seed=5)
    t0 = bo.totals(U)
    for _ in range(5):
        bo.step(U, bo.compute_dt(U, 0.3))
    t1 = bo.totals(U)
    assert t1[0] == pytest.approx(t0[0], rel=1e-13)
    assert t1[4] == pytest.approx(t0[4], rel=1e-13)


@pytest.mark.parametrize("maker", [two_level_2d, two_level_3d])
def test_subcycled_reflux_conserves_to_roundoff(O, maker):
    """Subcycling breaks conservation at coarse↔fine faces unless the flux
    register is applied (src/reflux.jl:91-93); a closed box must conserve."""
    m = maker()
    bo = O.BlockOracle(m)
    U = BM.smooth_block_state(m, seed=6)
    t0 = bo.totals(U)
    t, k = bo.evolve(U, 0.02, 0.3, subcycle=True)
    assert k > 0
    t1 = bo.totals(U)
    assert t1[0] == pytest.approx(t0[0], rel=1e-13)
    assert t1[4] == pytest.approx(t0[4], rel=1e-13)


def test_free_stream_preserved_across_hanging_faces(O):
    """A uniform moving state with outflow boundaries is a steady solution; every
    face carries F(W), so each leaf's net flux cancels to round-off (coarse faces
    see 2^(rank-1) fine sub-faces of a quarter of the area)."""
    for m in (two_level_2d(bc=1), two_level_3d(bc=1)):
        bo = O.BlockOracle(m)
        W = np.array([1.3, 0.2, -0.1, 0.05 if m.rank == 3 else 0.0, 0.7])
        r, vx, vy, vz, p = W
        u = np.array([r, r * vx, r * vy, r * vz, p / 0.4 + 0.5 * r * (vx * vx + vy * vy + vz * vz)])
        U = np.repeat(u[:, None], m.ncell, axis=1).copy()
        U0 = U.copy()
        for _ in range(3):
            bo.step(U, 1e-3)
        np.testing.assert_allclose(U, U0, rtol=1e-14, atol=1e-15)


def _sod_blocks():
    h0 = 1.0 / 64
    # coarse [0, .25) | fine [.25, .75) | coarse [.75, 1)
    return BM.BlockMesh(1, 1, (h0,), 1.4, [0, 1, 0], [[16, 1, 1], [64, 1, 1], [16, 1, 1]],
                        [[0, 0, 0], [32, 0, 0], [96, 0, 0]], [[1] * 6] * 3)


def _sod_state(m):
    U = np.empty((5, m.ncell))
    off = m.offsets
    for b in range(m.nblocks):
        X = m.cell_centers(b)[0].ravel()
        r = np.where(X < 0.5, 1.0, 0.125)
        p = np.where(X < 0.5, 1.0, 0.1)
        U[:, off[b]:off[b + 1]] = np.stack([r, 0 * r, 0 * r, 0 * r, p / (1.4 - 1.0)])
    return U


@pytest.mark.parametrize("subcycle", [False, True])
def test_refined_sod(O, subcycle):
    """test/test_amr.jl / test_reflux.jl analogue: refined Sod conserves mass and
    energy within rtol 1e-9 (outflow ends) and tracks the exact solution."""
    m = _sod_blocks()
    bo = O.BlockOracle(m)
    U = _sod_state(m)
    t0 = bo.totals(U)
    t, k = bo.evolve(U, 0.2, 0.4, subcycle=subcycle)
    assert t == pytest.approx(0.2, abs=1e-14)
    t1 = bo.totals(U)
    assert t1[0] == pytest.approx(t0[0], rel=1e-9)
    assert t1[4] == pytest.approx(t0[4], rel=1e-9)
    err = vol = 0.0
    for b in range(m.nblocks):
        X = m.cell_centers(b)[0].ravel()
        w = m.width(b)[0]
        rho = U[0, m.offsets[b]:m.offsets[b + 1]]
        ex = np.array([O.exact_riemann_sample((1, 0, 1), (0.125, 0, 0.1), 1.4, (x - 0.5) / t)[0] for x in X])
        err += np.sum(np.abs(rho - ex)) * w
        vol += w * len(X)
    assert err / vol < 0.01


def test_random_block_pairs_validate_and_conserve(O):
    m = BM.random_block_pairs(12, seed=7, rank=3, edges=(8, 16))
    m.validate()
    m.bc[:] = 2                                # close every free face
    bo = O.BlockOracle(m)
    assert bo.validate() == 0
    assert (m.level == 1).any()                # at least one 2:1 pair
    U = BM.smooth_block_state(m, seed=8)
    t0 = bo.totals(U)
    for _ in range(2):
        bo.step(U, bo.compute_dt(U, 0.3))
    t1 = bo.totals(U)
    assert t1[0] == pytest.approx(t0[0], rel=1e-13)
    assert t1[4] == pytest.approx(t0[4], rel=1e-13)


def test_validate_rejects_bad_meshes():
    with pytest.raises(ValueError, match="overlap"):
        BM.BlockMesh(2, 0, (0.1, 0.1), 1.4, [0, 0], [[4, 4, 1]] * 2, [[0, 0, 0], [2, 0, 0]],
                     [[1] * 6] * 2).validate()
    with pytest.raises(ValueError, match="level jump"):
        BM.BlockMesh(2, 2, (0.1, 0.1), 1.4, [0, 2], [[4, 4, 1]] * 2, [[0, 0, 0], [16, 0, 0]],
                     [[1] * 6] * 2).validate()
    with pytest.raises(ValueError, match="aligned"):
        BM.BlockMesh(2, 1, (0.1, 0.1), 1.4, [0, 1], [[4, 4, 1], [4, 4, 1]], [[0, 0, 0], [8, 1, 0]],
                     [[1] * 6] * 2).validate()
