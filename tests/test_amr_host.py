"""Host logic of the dynamic block AMR driver (no device): tagging from the
per-block indicator maxima, refinement with 2:1 balance, the source-block table
of the remap, and the block mesh it hands to the library's topology builder.
The indicator and the remap themselves come from the oracle here (the device
computes them in a run)."""
import numpy as np

from conftest import load_vespa

V = load_vespa()
from vespa_jl_amd import amr, blockmesh as BM  # noqa: E402


def _bump(qb, width):
    x, y = qb.cell_centers()
    U = amr.sedov_quadblocks_state(qb)
    U[0] = 1.0 + 0.5 * np.exp(-(x * x + y * y) / width)
    return U


def _regrid(O, qb, U, thr):
    old, old_mesh = list(qb.blocks), qb.mesh()
    bmax, _ = O.BlockOracle(old_mesh).indicator(U)
    qb.refine(qb.tags_from_indicator(bmax, thr, symmetric=True))
    return O.blocks_remap(qb.mesh(), old_mesh, qb.source_blocks(old), U)


def test_refine_balance_and_conservative_remap(O):
    qb = amr.QuadBlocks(6, 8, 3)
    U = _bump(qb, 0.01)
    tot0 = (U * qb.volumes()).sum(axis=1)
    for _ in range(4):
        U = _regrid(O, qb, U, 0.02)
        # injection keeps every conserved total (to round-off of the sums)
        np.testing.assert_allclose((U * qb.volumes()).sum(axis=1), tot0, rtol=1e-14, atol=1e-16)
        m = qb.mesh()
        m.validate()                       # overlap / 2:1 / alignment
        c = BM.describe(m)
        assert c["cells"] == len(qb.blocks) * 64
    levels = sorted({b[0] for b in qb.blocks})
    assert levels == [0, 1, 2, 3]
    # the blocks tile the square exactly
    area = sum(qb.h(l) ** 2 * 64 for (l, _, _) in qb.blocks)
    assert abs(area - 1.0) < 1e-12


def test_mesh_is_four_fold_symmetric(O):
    qb = amr.QuadBlocks(6, 8, 3)
    U = _bump(qb, 0.02)
    for _ in range(3):
        U = _regrid(O, qb, U, 0.02)
    n = {}
    for (l, ix, iy) in qb.blocks:
        nl = 6 << l
        n[(l, ix, iy)] = 1
        for img in ((l, nl - 1 - ix, iy), (l, ix, nl - 1 - iy), (l, iy, ix)):
            assert img in set(qb.blocks), img


def test_tags_and_source_blocks():
    """Only blocks below maxlevel above the threshold are tagged; mirror images
    join with symmetric=True; each child's source is its parent."""
    qb = amr.QuadBlocks(2, 4, 1)
    bm = np.zeros(4)
    bm[0] = 1.0                                   # block (0, 0, 0)
    assert qb.tags_from_indicator(bm, 0.5) == {(0, 0, 0)}
    assert qb.tags_from_indicator(bm, 0.5, symmetric=True) == set(qb.blocks)
    old = list(qb.blocks)
    qb.refine({(0, 0, 0)})
    src = qb.source_blocks(old)
    for q, (l, ix, iy) in enumerate(qb.blocks):
        assert old[src[q]] == ((l, ix, iy) if l == 0 else (0, ix >> 1, iy >> 1))
    assert qb.tags_from_indicator(np.ones(len(qb.blocks)), 0.5) == {b for b in qb.blocks if b[0] == 0}


def test_sedov_initial_state_deposits_the_energy():
    qb = amr.QuadBlocks(6, 8, 0)
    U = amr.sedov_quadblocks_state(qb)
    vol = qb.volumes()
    x, y = qb.cell_centers()
    hot = np.hypot(x, y) < 0.02
    assert hot.sum() == 4                                    # base 48^2: the four central cells
    E = (U[4] * vol).sum()
    assert abs(E - (1.0 + 1e-5 / 0.4 * (1.0 - vol[hot].sum()))) < 1e-12


def test_shock_radius_of_a_ring():
    N = 384
    c = -0.5 + (np.arange(N) + 0.5) / N
    X, Y = np.meshgrid(c, c)
    rho = 1.0 + np.exp(-((np.hypot(X, Y) - 0.2) / 0.01) ** 2)
    assert abs(amr.shock_radius(rho) - 0.2) < 2.0 / N


def _sweep_refine(nroot, maxlevel, leaves, tags):
    """2:1 balance by sweeps: split the tags, then every leaf that has a face
    neighbour more than one level finer (found through the finest-lattice
    positions along its faces), until none does. Slow, and the definition the
    work-list form in QuadBlocks.refine has to reproduce."""
    leaves = set(leaves)
    todo = set(tags)
    nmax = nroot << maxlevel

    def leaf_at(fx, fy):
        for l in range(maxlevel, -1, -1):
            s = 1 << (maxlevel - l)
            if (l, fx // s, fy // s) in leaves:
                return (l, fx // s, fy // s)
        return None

    while todo:
        for (l, ix, iy) in todo:
            if (l, ix, iy) in leaves:
                leaves.discard((l, ix, iy))
                leaves.update((l + 1, 2 * ix + cx, 2 * iy + cy) for cx in (0, 1) for cy in (0, 1))
        todo = set()
        for (l, ix, iy) in leaves:
            s = 1 << (maxlevel - l)
            edge = [(ix * s - 1, iy * s + t) for t in range(s)] + [(ix * s + s, iy * s + t) for t in range(s)] + \
                   [(ix * s + t, iy * s - 1) for t in range(s)] + [(ix * s + t, iy * s + s) for t in range(s)]
            for (px, py) in edge:
                if 0 <= px < nmax and 0 <= py < nmax:
                    nb = leaf_at(px, py)
                    if nb is not None and nb[0] > l + 1:
                        todo.add((l, ix, iy))
                        break
    return sorted(leaves, key=lambda b: (b[0], b[2], b[1]))


def test_refine_matches_sweep_balance():
    """The work-list balance gives the same leaves as splitting by sweeps, on
    random tag sequences over small quadtrees (1-4 roots per axis, up to 4
    levels); each split counts once."""
    rng = np.random.default_rng(7)
    for _ in range(60):
        nroot, ml = int(rng.integers(1, 5)), int(rng.integers(1, 5))
        qb = amr.QuadBlocks(nroot, 4, ml)
        for _ in range(int(rng.integers(1, 5))):
            cand = [b for b in qb.blocks if b[0] < ml]
            if not cand:
                break
            pick = rng.choice(len(cand), size=int(rng.integers(1, min(len(cand), 5) + 1)), replace=False)
            tags = {cand[i] for i in pick}
            want = _sweep_refine(nroot, ml, qb.blocks, tags)
            n0 = len(qb.blocks)
            nsplit = qb.refine(tags)
            assert qb.blocks == want
            assert len(qb.blocks) == n0 + 3 * nsplit
