"""Host topology of the batched block path (no device): the library's table
builder against an independent count of contact faces, and its mesh checks.

A link face is a face between two different blocks; there is one per finer-side
cell on every contact (HGBackend emits coarse-fine faces once per fine leaf,
lib/HGBackend/src/HGBackend.jl:202-276). Flux-register cells are the coarse
cells with finer neighbours (src/reflux.jl:51-103).
"""
import itertools

import numpy as np
import pytest

from conftest import load_vespa

V = load_vespa()
from vespa_jl_amd import blockmesh as BM  # noqa: E402
from vespa_jl_amd._lib import VespaHIPError  # noqa: E402


def contact_sets(m):
    """Brute force over block pairs: (#link faces, cells touching a link face,
    coarse cells with a finer neighbour)."""
    r = m.rank
    span = np.array([m.span(b) for b in range(m.nblocks)])
    lo = m.origin[:, :r]
    hi = lo + m.dims[:, :r].astype(np.int64) * span[:, None]
    faces = 0
    touch, reg = set(), set()

    def face_cells(b, d, side, olo, ohi):
        s = span[b]
        ranges = []
        for e in range(r):
            if e == d:
                ranges.append([m.dims[b, d] - 1 if side else 0])
            else:
                ranges.append(range((olo[e] - lo[b][e]) // s, (ohi[e] - lo[b][e]) // s))
        off = m.offsets[b]
        nx, ny = m.dims[b, 0], m.dims[b, 1]
        for ijk in itertools.product(*ranges):
            ijk = list(ijk) + [0] * (3 - r)
            yield int(off + (ijk[2] * ny + ijk[1]) * nx + ijk[0])

    for a in range(m.nblocks):
        for b in range(a + 1, m.nblocks):
            ov = np.minimum(hi[a], hi[b]) - np.maximum(lo[a], lo[b])
            if not ((ov >= 0).all() and (ov == 0).sum() == 1):
                continue
            d = int(np.nonzero(ov == 0)[0][0])
            olo = np.maximum(lo[a], lo[b])
            ohi = np.minimum(hi[a], hi[b])
            fine = a if m.level[a] >= m.level[b] else b
            nf = 1
            for e in range(r):
                if e != d:
                    nf *= (ohi[e] - olo[e]) // span[fine]
            faces += int(nf)
            a_hi = hi[a][d] == lo[b][d]   # a on the lo side
            touch.update(face_cells(a, d, 1 if a_hi else 0, olo, ohi))
            touch.update(face_cells(b, d, 0 if a_hi else 1, olo, ohi))
            if m.level[a] != m.level[b]:
                c = b if fine == a else a
                c_side = (1 if a_hi else 0) if c == a else (0 if a_hi else 1)
                reg.update(face_cells(c, d, c_side, olo, ohi))
    return faces, touch, reg


MESHES = {
    "2d_two_level": lambda: BM.BlockMesh(2, 1, (1 / 16, 1 / 16), 1.4, [0, 1, 1], [[8, 8, 1], [8, 8, 1], [8, 4, 1]],
                                         [[0, 0, 0], [16, 4, 0], [16, 0, 0]], [[2] * 4 + [1, 1]] * 3),
    "3d_two_level": lambda: BM.BlockMesh(3, 1, (1 / 8,) * 3, 1.4, [0, 1, 0], [[4, 4, 4], [8, 8, 8], [4, 4, 4]],
                                         [[0, 0, 0], [8, 0, 0], [16, 0, 0]], [[2] * 6] * 3),
    "1d_sod": lambda: BM.BlockMesh(1, 1, (1 / 64,), 1.4, [0, 1, 0], [[16, 1, 1], [64, 1, 1], [16, 1, 1]],
                                   [[0, 0, 0], [32, 0, 0], [96, 0, 0]], [[1] * 6] * 3),
    "random_40": lambda: BM.random_block_pairs(40, seed=7, rank=3, edges=(8, 16)),
    "random_2d": lambda: BM.random_block_pairs(30, seed=11, rank=2, edges=(8, 16, 24, 32)),
}


@pytest.mark.parametrize("name", sorted(MESHES))
def test_describe_matches_brute_force_contacts(name):
    m = MESHES[name]()
    m.validate()
    c = BM.describe(m)
    faces, touch, reg = contact_sets(m)
    assert c["cells"] == m.ncell
    assert c["link_faces"] == faces
    assert c["link_cells"] == len(touch)
    assert c["register_cells"] == len(reg)
    assert c["slot_faces"] == 2 * faces          # each face is listed by both of its cells
    assert c["tiles"] >= m.nblocks


def test_describe_tiles_follow_zchunk():
    m = BM.uniform_as_blocks((32, 24, 40), [(0, 1)] * 3, [(0, 0)] * 3, 1.4)
    whole = BM.describe(m)["tiles"]
    # x: 32 (one tile), y: 24 in tiles of 8 -> 3, z whole
    assert whole == 3
    assert BM.describe(m, zchunk=8)["tiles"] == 3 * 5
    assert BM.describe(m, zchunk=16)["tiles"] == 3 * 3    # 40 planes in 3 balanced chunks


@pytest.mark.parametrize("bad,match", [
    ("jump", "more than one level"),
    ("periodic_contact", "periodic"),
    ("misaligned", "aligned"),
])
def test_describe_rejects_bad_meshes(bad, match):
    if bad == "jump":
        m = BM.BlockMesh(2, 2, (0.1, 0.1), 1.4, [0, 2], [[4, 4, 1]] * 2, [[0, 0, 0], [16, 0, 0]], [[1] * 6] * 2)
    elif bad == "periodic_contact":
        m = BM.BlockMesh(2, 0, (0.1, 0.1), 1.4, [0, 0], [[4, 4, 1]] * 2, [[0, 0, 0], [4, 0, 0]],
                         [[0, 0, 1, 1, 1, 1]] * 2)
    else:
        m = BM.BlockMesh(2, 1, (0.1, 0.1), 1.4, [0, 1], [[4, 4, 1], [4, 4, 1]], [[0, 0, 0], [8, 1, 0]],
                         [[1] * 6] * 2)
    with pytest.raises(VespaHIPError, match=match):
        BM.describe(m)


def _ring_quadtree(rounds):
    """A quadtree of 6x6 root blocks (8x8 cells, 3 levels) refined around growing rings:
    deterministic meshes of 48 .. ~1200 blocks with every kind of coarse/fine face."""
    from vespa_jl_amd import amr
    qb = amr.QuadBlocks(6, 8, 3)
    for r in range(rounds):
        x, y = qb.cell_centers()
        rr = np.sqrt(x * x + y * y)
        g = np.zeros(len(qb.blocks))
        for q in range(len(qb.blocks)):
            s = rr[q * 64:(q + 1) * 64]
            g[q] = 1.0 if ((s > 0.05 + 0.03 * r) & (s < 0.15 + 0.03 * r)).any() else 0.0
        qb.refine(qb.tags_from_indicator(g, 0.5, symmetric=True))
    return qb.mesh()


@pytest.mark.parametrize("rounds,counts", [
    (3, {"cells": 21504, "tiles": 336, "link_faces": 5632, "link_cells": 9260, "register_cells": 692,
         "slot_faces": 11264}),
    (8, {"cells": 74496, "tiles": 1164, "link_faces": 19008, "link_cells": 32348, "register_cells": 1052,
         "slot_faces": 38016}),
])
def test_describe_counts_of_refined_quadtrees(rounds, counts):
    """Table sizes of the host builder on refined quadtrees, as first recorded with
    node-based maps (the builder's lookups changed, its tables must not)."""
    m = _ring_quadtree(rounds)
    assert BM.describe(m) == counts


@pytest.mark.parametrize("name,zchunk,digest", [
    ("ring3", 0, 0xf9056b64bf2defc7),
    ("ring8", 0, 0xff84b41d15663f51),
    ("pairs300_3d", 0, 0x82220695a4d1ae80),
    ("pairs300_3d", 8, 0xbe8548bde1bc6ebc),
    ("pairs1000_2d", 0, 0xf95618b0ccd893a3),
])
def test_threaded_builder_tables_equal_serial_ones(name, zchunk, digest):
    """The builder splits its work over host threads (contiguous block chunks,
    joined in order). Every derived table must be the serial walk's: the digests
    were recorded with the earlier single-threaded builder (hash-map lookups),
    and every thread count reproduces them. Exception: the 3D cost-sorted tile
    lists are grouped by tile shape class since the stage kernel has a
    compile-time form per shape; that moved the pairs300_3d/zchunk 0 digest
    (from 0xd9a2bf2a2e96a910) and left the other four, whose lists hold one class
    (2D) or were already grouped (zchunk 8), as recorded."""
    m = {"ring3": lambda: _ring_quadtree(3), "ring8": lambda: _ring_quadtree(8),
         "pairs300_3d": lambda: BM.random_block_pairs(300, seed=1),
         "pairs1000_2d": lambda: BM.random_block_pairs(1000, seed=3, rank=2)}[name]()
    for nt in (1, 2, 3, 8, 0):
        assert BM.table_digest(m, zchunk, nt) == digest, nt


def test_table_digest_sees_a_changed_table():
    """The digest covers the tile lists (zchunk) and the face walk (one block moved)."""
    m = BM.random_block_pairs(40, seed=2)
    assert BM.table_digest(m, 0) != BM.table_digest(m, 8)
    q = _ring_quadtree(1)
    q2 = BM.BlockMesh(q.rank, q.maxlevel, q.h0[:q.rank], q.gamma, q.level[::-1].copy(), q.dims[::-1].copy(),
                      q.origin[::-1].copy(), q.bc[::-1].copy())
    assert BM.describe(q2) == BM.describe(q) and BM.table_digest(q2) != BM.table_digest(q)
