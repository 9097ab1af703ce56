"""Dynamic block AMR driver on top of the batched block path (2D quadtree of
fixed-size blocks).

The reference refines leaves with a density-gradient indicator every `every`
steps and once before the first step, refine-only, with a conservative remap
(src/driver.jl:392-462; HGBackend refine + remap, lib/HGBackend/src/HGBackend.jl;
injection / volume mean, lib/RefMesh/src/RefMesh.jl:250-302). Here the unit of
refinement is a block of B×B cells: a block below maxlevel is refined (split
into four children of B×B cells at the next level) when any of its leaves has
a density_gradient_indicator (src/driver.jl:400-414) above `refine_above`, then
2:1 balance is restored across block faces.

The state never leaves the device during a run: the indicator is evaluated by
a kernel that returns one maximum per block, the host turns those into tags
and a new block list (the tree is a few thousand entries), and the device
context regrids in place: the library's threaded builder makes the new mesh's
tables, uploads them in one copy next to the current ones, and a remap kernel
injects each parent cell's conserved values into its children device to
device, so Σ U·V is unchanged; no device memory is allocated unless the mesh
outgrows every earlier one. Between
regrids the leaves are advanced by the batched kernel (one persistent launch
per RK stage).
"""
from __future__ import annotations

import numpy as np

import time

from .blockmesh import BlockMesh, BlockSimulation, REF


class QuadBlocks:
    """Blocks of B×B cells on the square [x0, x0 + L]², level 0 = nroot×nroot
    root blocks; a level-l block (l ≤ maxlevel) has index (ix, iy) in units of its
    own size."""

    def __init__(self, nroot: int, B: int, maxlevel: int, x0: float = -0.5, L: float = 1.0, bc: int = REF,
                 gamma: float = 1.4):
        self.nroot, self.B, self.maxlevel = int(nroot), int(B), int(maxlevel)
        self.x0, self.L, self.bc, self.gamma = float(x0), float(L), int(bc), float(gamma)
        self.blocks = [(0, ix, iy) for iy in range(nroot) for ix in range(nroot)]

    # -- geometry
    def h(self, level: int) -> float:
        return self.L / (self.nroot * self.B * (1 << level))

    def mesh(self) -> BlockMesh:
        ml = self.maxlevel
        a = np.array(self.blocks, dtype=np.int64).reshape(-1, 3)
        lv, ix, iy = a[:, 0], a[:, 1], a[:, 2]
        n_l = np.int64(self.nroot) << lv                  # blocks per axis at each block's level
        assert ((0 <= ix) & (ix < n_l) & (0 <= iy) & (iy < n_l)).all()
        span = np.int64(self.B) << (ml - lv)              # lattice units per block edge
        org = np.zeros((len(a), 3), dtype=np.int64)
        org[:, 0] = ix * span
        org[:, 1] = iy * span
        dims = np.tile(np.array([self.B, self.B, 1], dtype=np.int32), (len(a), 1))
        bc = np.tile(np.array([self.bc] * 4 + [1, 1], dtype=np.int32), (len(a), 1))
        h0 = self.L / (self.nroot * self.B)
        return BlockMesh(2, ml, (h0, h0), self.gamma, lv.astype(np.int32), dims, org, bc, x0=(self.x0, self.x0))

    def cell_centers(self):
        """(x, y) of every cell in leaf order (block by block, x fastest)."""
        xs, ys = [], []
        for (l, ix, iy) in self.blocks:
            h = self.h(l)
            i = np.arange(self.B)
            X, Y = np.meshgrid(self.x0 + (ix * self.B + i + 0.5) * h, self.x0 + (iy * self.B + i + 0.5) * h)
            xs.append(X.ravel())
            ys.append(Y.ravel())
        return np.concatenate(xs), np.concatenate(ys)

    def volumes(self):
        return np.concatenate([np.full(self.B * self.B, self.h(l) ** 2) for (l, _, _) in self.blocks])

    # -- rasterisation on the finest uniform grid (injection)
    def raster(self, Uf: np.ndarray) -> np.ndarray:
        """Fields (nf, ncell) → (nf, N, N) on the finest grid, N = nroot·B·2^maxlevel."""
        N = self.nroot * self.B << self.maxlevel
        out = np.empty((Uf.shape[0], N, N), dtype=Uf.dtype)
        B2 = self.B * self.B
        for q, (l, ix, iy) in enumerate(self.blocks):
            r = 1 << (self.maxlevel - l)
            blk = Uf[:, q * B2:(q + 1) * B2].reshape(-1, self.B, self.B)
            fine = np.repeat(np.repeat(blk, r, axis=1), r, axis=2)
            y0, x0 = iy * self.B * r, ix * self.B * r
            out[:, y0:y0 + self.B * r, x0:x0 + self.B * r] = fine
        return out

    # -- tagging (regrid!, src/driver.jl:440-452)
    def tags_from_indicator(self, block_max, refine_above: float, symmetric: bool = False) -> set:
        """Blocks below maxlevel whose largest leaf indicator exceeds refine_above.
        block_max: one value per block in the current order (the per-block
        maximum of density_gradient_indicator). symmetric: also tag the mirror
        images of every tagged block (x, y and diagonal mirrors of the square),
        so that round-off differences between symmetric cells cannot make the
        mesh asymmetric."""
        block_max = np.asarray(block_max)
        if block_max.shape != (len(self.blocks),):
            raise ValueError("one indicator value per block")
        tags = {b for b, g in zip(self.blocks, block_max) if b[0] < self.maxlevel and g > refine_above}
        if symmetric:
            leaves = set(self.blocks)
            for (l, ix, iy) in list(tags):
                nl = self.nroot << l
                for img in ((l, nl - 1 - ix, iy), (l, ix, nl - 1 - iy), (l, nl - 1 - ix, nl - 1 - iy),
                            (l, iy, ix), (l, nl - 1 - iy, ix), (l, iy, nl - 1 - ix), (l, nl - 1 - iy, nl - 1 - ix)):
                    if img in leaves:
                        tags.add(img)
        return tags

    def refine(self, tags: set) -> int:
        """Refine the tagged blocks, then restore 2:1 balance across faces.
        Returns the number of blocks split.

        Work-list form: a leaf at level l whose face neighbour region (the
        level-l slot across one of its faces, inside the square) lies in a leaf
        coarser than l - 1 forces that leaf to split; only leaves created by a
        split can break the balance, so each new leaf is checked once (and its
        own check repeats until the slot's leaf is fine enough). Every split is
        forced, so the result is the smallest balanced refinement that splits
        the tags, independent of the order of the work list."""
        leaves = set(self.blocks)
        queue = []

        def split(b):
            l, ix, iy = b
            leaves.discard(b)
            for cy in (0, 1):
                for cx in (0, 1):
                    c = (l + 1, 2 * ix + cx, 2 * iy + cy)
                    leaves.add(c)
                    queue.append(c)

        nsplit = 0
        for b in sorted(tags):
            if b in leaves:
                split(b)
                nsplit += 1
        while queue:
            c = queue.pop()
            if c not in leaves:
                continue                        # split since: its children are queued
            lc, cx, cy = c
            if lc < 2:
                continue
            n = self.nroot << lc
            for px, py in ((cx - 1, cy), (cx + 1, cy), (cx, cy - 1), (cx, cy + 1)):
                if not (0 <= px < n and 0 <= py < n):
                    continue
                while True:
                    # the leaf holding slot (lc, px, py): itself or an ancestor
                    for up in range(lc + 1):
                        a = (lc - up, px >> up, py >> up)
                        if a in leaves:
                            break
                    else:
                        break                   # the slot is split finer than lc
                    if up < 2:
                        break
                    split(a)
                    nsplit += 1
        # order: by level, then row-major, for a deterministic layout
        self.blocks = sorted(leaves, key=lambda b: (b[0], b[2], b[1]))
        return nsplit

    def source_blocks(self, old_blocks: list) -> np.ndarray:
        """For every current block, the index in old_blocks of the same block or
        of the ancestor it was split from: the table the remap injects by."""
        where = {b: q for q, b in enumerate(old_blocks)}
        out = np.empty(len(self.blocks), dtype=np.int32)
        for q, b in enumerate(self.blocks):
            s = where.get(b)
            if s is None:
                l, ix, iy = b
                for up in range(1, l + 1):
                    s = where.get((l - up, ix >> up, iy >> up))
                    if s is not None:
                        break
                else:
                    raise ValueError(f"block {b} has no ancestor in the old mesh")
            out[q] = s
        return out


def sedov_quadblocks_state(qb: QuadBlocks, E0: float = 1.0, rho0: float = 1.0, p0: float = 1e-5,
                           r_hot: float = 0.02, dtype=np.float64) -> np.ndarray:
    """Sedov–Taylor initial state on the current blocks: ρ = rho0, v = 0,
    p = p0 outside, and the energy E0 deposited uniformly (as pressure) in the
    cells whose centre lies within r_hot of the origin."""
    x, y = qb.cell_centers()
    vol = qb.volumes()
    hot = np.hypot(x, y) < r_hot
    if not hot.any():
        raise ValueError("no cell centre within r_hot")
    e_hot = E0 / vol[hot].sum()                      # energy density in the hot cells
    g = qb.gamma
    E = np.where(hot, e_hot, p0 / (g - 1.0))
    n = x.size
    U = np.stack([np.full(n, rho0), np.zeros(n), np.zeros(n), np.zeros(n), E])
    return U.astype(dtype)


def evolve_amr(qb: QuadBlocks, U0: np.ndarray, tfinal: float, cfl: float = 0.4, refine_above: float = 0.1,
               every: int = 8, dtype=np.float64, mode: str = "exact", snapshots=(), max_steps: int = 100000,
               symmetric: bool = False, graphs: bool = False):
    """evolve! with a RefinementPolicy (src/driver.jl:483-508) on the quadtree
    `qb` from the conserved state U0 (5, ncell) in its leaf order: regrid before
    the first step and every `every` steps (refine-only), leaves advanced by the
    batched kernel, the state on the device from the first upload to the final
    download. Returns a dict with the final blocks/state, step count, leaf
    counts, the raster density at the requested snapshot times and the wall
    time spent in the regrid phases (ms): indicator kernel + per-block maxima to
    the host, host tagging and refinement (tree edit and the new block table),
    and the in-place regrid of the device context (tables built by the host
    builder's threads, uploaded in one copy, and the device remap, all under
    `tables`)."""
    totals0 = (np.asarray(U0, dtype=np.float64) * qb.volumes()).sum(axis=1)
    sim = BlockSimulation(qb.mesh(), dtype, cfl=cfl)
    sim.set_mode(mode)
    if not graphs:
        # Steps launched eagerly: with a regrid check every `every` steps a step
        # graph would be captured, instantiated and destroyed per regrid for a
        # few replays (measured on the Sedov run: whole run 0.069 -> 0.059 s,
        # regrid tables 6.5 -> 5.3 ms without graphs).
        sim.set_launch(zchunk=sim.info()["zchunk"], use_graph=False)
    sim.upload(np.ascontiguousarray(U0, dtype=dtype))
    t, steps = 0.0, 0
    snaps = {}
    pending = sorted(snapshots)
    leaves_hist = []
    phase = {"indicator": 0.0, "refine": 0.0, "tables": 0.0}
    nregrid = ncheck = 0
    B = qb.B
    try:
        while t < tfinal * (1 - 1e-12) and steps < max_steps:
            t0 = time.perf_counter()
            bmax = sim.indicator()
            t1 = time.perf_counter()
            tags = qb.tags_from_indicator(bmax, refine_above, symmetric=symmetric)
            phase["indicator"] += t1 - t0
            ncheck += 1
            if tags:
                old = list(qb.blocks)
                qb.refine(tags)
                src = qb.source_blocks(old)
                mesh = qb.mesh()
                t2 = time.perf_counter()
                sim.regrid(mesh, src)   # tables, their upload and the remap, in place
                t3 = time.perf_counter()
                phase["refine"] += t2 - t1
                phase["tables"] += t3 - t2
                nregrid += 1
            sim.t = t
            stop = tfinal
            if pending and pending[0] < tfinal:
                stop = pending[0]
            tt, k = sim.evolve(tfinal=stop, max_steps=every)
            t = tt
            steps += k
            leaves_hist.append(len(qb.blocks) * B * B)
            while pending and t >= pending[0] * (1 - 1e-12):
                rho = sim.state()[:1].astype(np.float64)
                snaps[pending.pop(0)] = (t, qb.raster(rho)[0])
        U = sim.state()
    finally:
        sim.close()
    vol = qb.volumes()
    totals = (U.astype(np.float64) * vol).sum(axis=1)
    return {"blocks": list(qb.blocks), "qb": qb, "U": U, "t": t, "steps": steps, "leaves": len(qb.blocks) * B * B,
            "max_leaves": max(leaves_hist) if leaves_hist else 0, "totals0": totals0, "totals": totals,
            "snapshots": snaps, "indicator_calls": ncheck, "regrids": nregrid,
            "regrid_ms": {k: 1e3 * v for k, v in phase.items()}}


def run_sedov(nroot: int = 6, B: int = 8, maxlevel: int = 3, tfinal: float = 0.05, cfl: float = 0.4,
              refine_above: float = 0.1, every: int = 8, dtype=np.float64, mode: str = "exact",
              snapshots=(), max_steps: int = 100000, symmetric: bool = True, graphs: bool = False):
    """Sedov blast on [-0.5, 0.5]² with base (nroot·B)² cells and up to maxlevel
    levels of refinement, reflecting walls, through evolve_amr."""
    qb = QuadBlocks(nroot, B, maxlevel)
    U0 = sedov_quadblocks_state(qb, dtype=np.float64).astype(dtype)
    return evolve_amr(qb, U0, tfinal, cfl=cfl, refine_above=refine_above, every=every, dtype=dtype, mode=mode,
                      snapshots=snapshots, max_steps=max_steps, symmetric=symmetric, graphs=graphs)


def oblique_sod_state(qb: QuadBlocks, angle: float = np.pi / 4, gamma: float = 1.4, dtype=np.float64) -> np.ndarray:
    """Sod's Riemann problem with its diaphragm through the centre of the square
    along the normal n = (cos angle, sin angle): left state (ρ 1, p 1) where
    n·x < 0, right state (ρ 0.125, p 0.1), at rest (the planar problem of
    src/problem.jl:61-71 turned by `angle`). Its exact solution is the 1D one in
    ξ = n·x until boundary effects arrive."""
    x, y = qb.cell_centers()
    xi = np.cos(angle) * x + np.sin(angle) * y
    left = xi < 0
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    z = np.zeros_like(rho)
    return np.stack([rho, z, z, z, p / (gamma - 1.0)]).astype(dtype)


def shock_radius(rho_raster: np.ndarray, x0: float = -0.5, L: float = 1.0) -> float:
    """Radius of the density maximum, averaged over angle (radial bins of one
    raster cell)."""
    N = rho_raster.shape[0]
    h = L / N
    c = x0 + (np.arange(N) + 0.5) * h
    X, Y = np.meshgrid(c, c)
    r = np.hypot(X, Y)
    bins = (r / h).astype(int)
    sums = np.bincount(bins.ravel(), weights=rho_raster.ravel())
    cnt = np.bincount(bins.ravel())
    prof = sums / np.maximum(cnt, 1)
    k = int(np.argmax(prof[: N // 2]))
    return (k + 0.5) * h


def shock_radius_mean(qb: QuadBlocks, rho: np.ndarray) -> float:
    """Area-weighted mean radius of the cells whose density lies above the
    midpoint between the background (1) and the peak (the estimator of the
    reference's Sedov test, test/test_sedov.jl:11-27)."""
    x, y = qb.cell_centers()
    w = qb.volumes()
    thresh = 0.5 * (1.0 + rho.max())
    sel = rho >= thresh
    return float((np.hypot(x, y)[sel] * w[sel]).sum() / w[sel].sum()) if sel.any() else 0.0


def quadrant_masses(qb: QuadBlocks, rho: np.ndarray) -> np.ndarray:
    """Mass in each quadrant about the centre (test/test_sedov.jl:58-69)."""
    x, y = qb.cell_centers()
    q = (x >= 0).astype(int) + 2 * (y >= 0).astype(int)
    return np.bincount(q, weights=rho * qb.volumes(), minlength=4)
