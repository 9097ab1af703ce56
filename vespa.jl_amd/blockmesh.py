"""Block-structured AMR mesh description (host data) for the batched block path.

A mesh is a list of blocks. Block b has a refinement level L_b (0 = base), a
cell count per axis ``dims[b]`` and an integer ``origin[b]`` on the lattice of the
finest cells (one lattice unit = one cell of level ``maxlevel``), so a level-L
cell spans 2^(maxlevel−L) lattice units and ``h0/2^L`` physical units. Leaves
(cells) are ordered block by block and column-major (x fastest) inside a block;
field f of cell c of block b lives at ``U[f*ncell + offset[b] + c]``. Blocks touch
only geometrically; ``bc[b]`` gives the boundary condition of each block face
where no other block touches it (periodic wraps the block onto itself).

Constraints (checked by ``validate``): adjacent leaves differ by at most one
level (2:1 balance, as HGBackend keeps it, lib/HGBackend/src/HGBackend.jl:26-29)
and a fine block's origin and extent are aligned to the cells of any coarser
block it touches, so each coarse face sees the 2^(rank−1) fine cells of one block.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

PER, OUT, REF = 0, 1, 2


@dataclass
class BlockMesh:
    rank: int
    maxlevel: int
    h0: tuple                      # level-0 cell width per axis
    gamma: float
    level: np.ndarray              # (nb,) int32
    dims: np.ndarray               # (nb, 3) int32 (axes ≥ rank hold 1)
    origin: np.ndarray             # (nb, 3) int64 lattice units
    bc: np.ndarray                 # (nb, 6) int32, [2*axis + side]
    x0: tuple = (0.0, 0.0, 0.0)    # physical coordinate of lattice point 0

    def __post_init__(self):
        self.level = np.ascontiguousarray(self.level, dtype=np.int32)
        self.dims = np.ascontiguousarray(self.dims, dtype=np.int32).reshape(-1, 3)
        self.origin = np.ascontiguousarray(self.origin, dtype=np.int64).reshape(-1, 3)
        self.bc = np.ascontiguousarray(self.bc, dtype=np.int32).reshape(-1, 6)
        self.h0 = tuple(float(h) for h in self.h0) + (1.0,) * (3 - len(self.h0))
        self.x0 = tuple(float(x) for x in self.x0) + (0.0,) * (3 - len(self.x0))

    @property
    def nblocks(self) -> int:
        return len(self.level)

    @property
    def counts(self) -> np.ndarray:
        return np.prod(self.dims.astype(np.int64), axis=1)

    @property
    def offsets(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)

    @property
    def ncell(self) -> int:
        return int(self.counts.sum())

    def span(self, b: int) -> int:
        return 1 << (self.maxlevel - int(self.level[b]))

    def width(self, b: int):
        return tuple(self.h0[d] / (1 << int(self.level[b])) for d in range(3))

    def cell_centers(self, b: int):
        """Physical centres of block b's cells as (z, y, x)-shaped arrays per axis."""
        s = self.span(b)
        u = self.h0[0] / (1 << self.maxlevel)
        out = []
        for d in range(3):
            n = int(self.dims[b, d])
            ud = self.h0[d] / (1 << self.maxlevel)
            lat = self.origin[b, d] + (np.arange(n) + 0.5) * s
            out.append(self.x0[d] + lat * ud)
        Z, Y, X = np.meshgrid(out[2], out[1], out[0], indexing="ij")
        return X, Y, Z

    def validate(self) -> None:
        """Raise ValueError when two blocks overlap, a level jump exceeds one or a
        fine block is not aligned to the coarse cells it touches."""
        nb, r = self.nblocks, self.rank
        lo = self.origin[:, :r]
        hi = lo + self.dims[:, :r].astype(np.int64) * np.array([self.span(b) for b in range(nb)])[:, None]
        for b in range(nb):
            if (self.dims[b, :r] < 1).any() or (self.dims[b, r:] != 1).any():
                raise ValueError(f"block {b}: bad dims {self.dims[b]}")
            if (self.origin[b, :r] % self.span(b)).any():
                raise ValueError(f"block {b}: origin not aligned to its own cells")
            for d in range(r):
                if (self.bc[b, 2 * d] == PER) != (self.bc[b, 2 * d + 1] == PER):
                    raise ValueError(f"block {b}: periodic must be set on both sides of axis {d}")
        for a in range(nb):
            for b in range(a + 1, nb):
                ov = np.minimum(hi[a], hi[b]) - np.maximum(lo[a], lo[b])
                if (ov > 0).all():
                    raise ValueError(f"blocks {a} and {b} overlap")
                touch = (ov >= 0).all() and (ov == 0).sum() == 1
                if not touch:
                    continue
                la, lb = int(self.level[a]), int(self.level[b])
                if abs(la - lb) > 1:
                    raise ValueError(f"blocks {a} and {b} touch across a level jump of {abs(la - lb)}")
                d = int(np.nonzero(ov == 0)[0][0])
                if self.bc[a, 2 * d] == PER or self.bc[b, 2 * d] == PER:
                    raise ValueError(f"blocks {a} and {b} touch along a periodic axis")
                if la != lb:
                    f, c = (a, b) if la > lb else (b, a)
                    sc = self.span(c)
                    for e in range(r):
                        if e == d:
                            continue
                        # fine extent must start and end on coarse cell boundaries inside the overlap
                        olo = max(lo[f][e], lo[c][e])
                        ohi = min(hi[f][e], hi[c][e])
                        if (olo - lo[c][e]) % sc or (ohi - lo[c][e]) % sc:
                            raise ValueError(f"fine block {f} is not aligned to the cells of block {c}")


def components(mesh: BlockMesh) -> list:
    """Groups of blocks connected by face contact (each sorted by block index).
    Blocks of different groups never exchange fluxes, so a group evolves the same
    inside the whole mesh and on its own."""
    nb, r = mesh.nblocks, mesh.rank
    span = np.array([mesh.span(b) for b in range(nb)], dtype=np.int64)
    lo = mesh.origin[:, :r]
    hi = lo + mesh.dims[:, :r].astype(np.int64) * span[:, None]
    parent = list(range(nb))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for a in range(nb):
        ov = np.minimum(hi[a], hi) - np.maximum(lo[a], lo)          # (nb, r)
        touch = (ov >= 0).all(axis=1) & ((ov == 0).sum(axis=1) == 1)
        touch[a] = False
        for b in np.nonzero(touch)[0]:
            ra, rb = find(a), find(int(b))
            if ra != rb:
                parent[max(ra, rb)] = min(ra, rb)
    groups = {}
    for b in range(nb):
        groups.setdefault(find(b), []).append(b)
    return [sorted(g) for g in groups.values()]


def submesh(mesh: BlockMesh, blocks) -> tuple:
    """The mesh made of the listed blocks (in the given order) and, for each of
    its cells, the index of the same cell in the parent's leaf order."""
    blocks = list(blocks)
    sub = BlockMesh(mesh.rank, mesh.maxlevel, mesh.h0[:mesh.rank], mesh.gamma, mesh.level[blocks],
                    mesh.dims[blocks], mesh.origin[blocks], mesh.bc[blocks], mesh.x0)
    off = mesh.offsets
    cells = np.concatenate([np.arange(off[b], off[b + 1]) for b in blocks])
    return sub, cells


def uniform_as_blocks(dims, domain, bcs, gamma) -> BlockMesh:
    """One block covering a UniformMesh (the RefMesh case)."""
    rank = len(dims)
    h0 = tuple((hi - lo) / n for (lo, hi), n in zip(domain, dims))
    d3 = list(dims) + [1] * (3 - rank)
    bc = []
    for d in range(3):
        if d < rank:
            lo, hi = bcs[d] if isinstance(bcs, (list, tuple)) else (bcs, bcs)
            bc += [lo, hi]
        else:
            bc += [OUT, OUT]
    return BlockMesh(rank, 0, h0, gamma, [0], [d3], [[0, 0, 0]], [bc],
                     x0=tuple(lo for lo, _ in domain))


def random_block_pairs(nblocks: int, seed: int, rank: int = 3, edges=(8, 16, 24, 32), p_link: float = 0.5,
                       p_fine: float = 0.5, gamma: float = 5.0 / 3.0) -> BlockMesh:
    """The batched-blocks workload: `nblocks` blocks with edges drawn uniformly from
    `edges`, laid out far apart as independent units, except that a random subset
    is glued face to face in pairs: conforming pairs (same level, same face extent)
    and 2:1 pairs (a level-1 block against a level-0 block, aligned to its cells).
    Faces that touch nothing get outflow or reflecting boundaries at random."""
    rng = np.random.default_rng(seed)
    maxlevel = 1
    level, dims, origin, bc = [], [], [], []
    slot = 0
    pitch = 4 * max(edges) * 2            # lattice units between independent units

    def place(lv, e, org):
        level.append(lv)
        d3 = [e if d < rank else 1 for d in range(3)]
        dims.append(d3)
        origin.append(org)
        bc.append([int(rng.integers(1, 3)) if d < rank else OUT for d in range(3) for _ in (0, 1)])

    b = 0
    while b < nblocks:
        base = [slot * pitch if d == 0 else 0 for d in range(3)]
        slot += 1
        e = int(rng.choice(edges))
        if b + 1 < nblocks and rng.random() < p_link:
            ax = int(rng.integers(0, rank))
            if rng.random() < p_fine:
                # coarse block (level 0, span 2) and a fine block (level 1, span 1)
                ef = int(rng.choice([x for x in edges if x // 2 <= e]))
                place(0, e, base)
                org = list(base)
                org[ax] = base[ax] + 2 * e            # touches the coarse hi face
                for d in range(rank):
                    if d != ax:
                        room = 2 * e - ef             # lattice units free along d
                        org[d] = base[d] + 2 * int(rng.integers(0, room // 2 + 1)) if room > 0 else base[d]
                place(1, ef, org)
            else:
                place(0, e, base)
                org = list(base)
                org[ax] = base[ax] + 2 * e
                place(0, e, org)
            b += 2
        else:
            place(0, e, base)
            b += 1
    h0 = tuple([1.0 / 32] * rank)
    return BlockMesh(rank, maxlevel, h0, gamma, level, dims, origin, bc)


def smooth_block_state(mesh: BlockMesh, seed: int, dtype=np.float64):
    """A smooth seeded state per block: ρ, p ∈ [0.5, 2], |v| ≤ 0.3 (conserved, SoA)."""
    rng = np.random.default_rng(seed)
    U = np.empty((5, mesh.ncell), dtype=dtype)
    off = mesh.offsets
    g = mesh.gamma
    for b in range(mesh.nblocks):
        X, Y, Z = mesh.cell_centers(b)
        k = rng.uniform(1.0, 4.0, (5, 3))
        ph = rng.uniform(0, 2 * np.pi, 5)
        f = [np.sin(2 * np.pi * (k[i, 0] * X + k[i, 1] * Y + k[i, 2] * Z) + ph[i]) for i in range(5)]
        rho = 1.25 + 0.75 * f[0]
        vx, vy, vz = 0.17 * f[1], 0.17 * f[2], 0.17 * f[3]
        if mesh.rank < 3:
            vz = 0.0 * vz
        if mesh.rank < 2:
            vy = 0.0 * vy
        p = 1.25 + 0.75 * f[4]
        W = np.stack([rho, vx, vy, vz, p]).reshape(5, -1).astype(dtype)
        T = np.dtype(dtype).type
        r, a, bb, c, pp = W
        E = pp / (T(g) - T(1)) + (T(0.5) * r) * ((a * a + bb * bb) + c * c)
        U[:, off[b]:off[b + 1]] = np.stack([r, r * a, r * bb, r * c, E])
    return U


# ---------------------------------------------------------------- device context

def _mesh_args(mesh: BlockMesh):
    import ctypes as C
    h0 = (C.c_double * 3)(*mesh.h0)
    lv = np.ascontiguousarray(mesh.level, dtype=np.int32)
    dm = np.ascontiguousarray(mesh.dims, dtype=np.int32)
    og = np.ascontiguousarray(mesh.origin, dtype=np.int64)
    bc = np.ascontiguousarray(mesh.bc, dtype=np.int32)
    i32p, i64p = C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    return (h0, lv.ctypes.data_as(i32p), dm.ctypes.data_as(i32p), og.ctypes.data_as(i64p), bc.ctypes.data_as(i32p),
            (lv, dm, og, bc))


COUNT_NAMES = ("cells", "tiles", "link_faces", "link_cells", "register_cells", "slot_faces")


def describe(mesh: BlockMesh, zchunk: int = 0) -> dict:
    """Validate a mesh with the library's topology builder (no device needed)
    and return the sizes of its derived tables."""
    import ctypes as C
    from ._lib import check, load_library
    L = load_library()
    h0, lv, dm, og, bc, keep = _mesh_args(mesh)
    out = (C.c_int64 * 6)()
    check(L.vh_blocks_describe(mesh.rank, mesh.nblocks, mesh.maxlevel, h0, lv, dm, og, bc, int(zchunk), out))
    return dict(zip(COUNT_NAMES, [int(v) for v in out]))


def table_digest(mesh: BlockMesh, zchunk: int = 0, nthreads: int = 0) -> int:
    """Digest of every table the library's builder derives for `mesh`, built
    with `nthreads` host threads (0: default, 1: serial); no device needed."""
    import ctypes as C
    from ._lib import check, load_library
    L = load_library()
    h0, lv, dm, og, bc, keep = _mesh_args(mesh)
    out = C.c_uint64()
    check(L.vh_blocks_table_digest(mesh.rank, mesh.nblocks, mesh.maxlevel, h0, lv, dm, og, bc, int(zchunk),
                                   int(nthreads), C.byref(out)))
    return int(out.value)


class BlockSimulation:
    """Device-resident simulation on a block mesh (the HGBackend Simulation,
    src/driver.jl:27-79): one persistent launch per RK stage over every block,
    link faces between blocks, flux registers for subcycling."""

    def __init__(self, mesh: BlockMesh, eltype=np.float64, cfl: float = 0.4):
        import ctypes as C
        from . import _lib
        from ._lib import check, load_library
        self._C, self._check = C, check
        self.mesh = mesh
        self.eltype = np.dtype(eltype)
        if self.eltype not in (np.float64, np.float32):
            raise ValueError(f"eltype must be float64 or float32, got {self.eltype}")
        self.cfl = float(cfl)
        self.t = 0.0
        self.step_count = 0
        self._lib = load_library()
        h0, lv, dm, og, bc, keep = _mesh_args(mesh)
        h = C.c_void_p()
        check(self._lib.vh_blocks_create(C.byref(h), mesh.rank, mesh.nblocks, mesh.maxlevel, h0,
                                         float(mesh.gamma), lv, dm, og, bc,
                                         _lib.VH_F64 if self.eltype == np.float64 else _lib.VH_F32))
        self._h = h
        self.ncell = int(self._lib.vh_blocks_ncell(h))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.vh_blocks_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, U: np.ndarray):
        U = np.ascontiguousarray(U, dtype=self.eltype)
        if U.size != 5 * self.ncell:
            raise ValueError(f"state must hold 5*{self.ncell} values")
        self._check(self._lib.vh_blocks_upload(self._h, U.ctypes.data_as(self._C.c_void_p)))

    def state(self) -> np.ndarray:
        out = np.empty((5, self.ncell), dtype=self.eltype)
        self._check(self._lib.vh_blocks_download(self._h, out.ctypes.data_as(self._C.c_void_p)))
        return out

    def set_mode(self, mode: str):
        from . import _lib
        self._check(self._lib.vh_blocks_set_mode(self._h, {"exact": _lib.VH_EXACT, "fast": _lib.VH_FAST}[mode]))

    # ---- self-gravity (src/gravity.jl) on the block mesh ----
    def enable_gravity(self, G: float = 1.0, maxiter: int = 500, tol: float | None = None,
                       project_mean: bool = True):
        """enable_gravity!(sim; G, maxiter, tol, project_mean) (src/gravity.jl:63-80):
        the composite Poisson operator of this mesh, φ = 0. tol defaults to 1e-8
        (Float64) / 1e-5 (Float32) as the reference's."""
        if tol is None:
            tol = 1e-8 if self.eltype == np.float64 else 1e-5
        self._check(self._lib.vh_blocks_enable_gravity(self._h, float(G), int(maxiter), float(tol),
                                                       int(bool(project_mean))))
        self.gravity_G = float(G)

    def disable_gravity(self):
        self._check(self._lib.vh_blocks_disable_gravity(self._h))
        self.gravity_G = None

    def solve_poisson(self) -> tuple:
        """solve_poisson! (src/gravity.jl:248-322) → (iterations, relative residual)."""
        C = self._C
        it, rr = C.c_int64(), C.c_double()
        self._check(self._lib.vh_blocks_solve_poisson(self._h, C.byref(it), C.byref(rr)))
        return int(it.value), float(rr.value)

    @property
    def phi(self) -> np.ndarray:
        out = np.empty(self.ncell, dtype=self.eltype)
        self._check(self._lib.vh_blocks_phi_download(self._h, out.ctypes.data_as(self._C.c_void_p)))
        return out

    @phi.setter
    def phi(self, value):
        v = np.ascontiguousarray(value, dtype=self.eltype)
        if v.size != self.ncell:
            raise ValueError(f"φ must hold {self.ncell} values")
        self._check(self._lib.vh_blocks_phi_upload(self._h, v.ctypes.data_as(self._C.c_void_p)))

    def apply_laplacian(self, x) -> np.ndarray:
        """A·x with A = −∇²·V, the composite operator of this mesh."""
        from ._lib import DeviceBuffer
        x = np.ascontiguousarray(x, dtype=self.eltype)
        bx, bo = DeviceBuffer(x.nbytes).upload(x), DeviceBuffer(x.nbytes)
        try:
            self._check(self._lib.vh_blocks_apply_laplacian(self._h, bx.ptr, bo.ptr))
            return bo.download(np.empty_like(x))
        finally:
            bx.free()
            bo.free()

    def gravity_accel(self) -> np.ndarray:
        """g = −∇φ of the current φ, shape (3, ncell)."""
        from ._lib import DeviceBuffer
        out = np.empty((3, self.ncell), dtype=self.eltype)
        b = DeviceBuffer(out.nbytes)
        try:
            self._check(self._lib.vh_blocks_gravity_accel(self._h, b.ptr))
            return b.download(out)
        finally:
            b.free()

    def set_launch(self, zchunk: int = 8, max_workgroups: int = 0, use_graph: bool = True):
        self._check(self._lib.vh_blocks_set_launch(self._h, int(zchunk), int(max_workgroups), int(bool(use_graph))))

    def set_tiling(self, max_tile_x: int = 32):
        self._check(self._lib.vh_blocks_set_tiling(self._h, int(max_tile_x)))

    def info(self) -> dict:
        C = self._C
        out = (C.c_int64 * 6)()
        zc, wg = C.c_int32(), C.c_int32()
        self._check(self._lib.vh_blocks_info(self._h, out, C.byref(zc), C.byref(wg)))
        d = dict(zip(COUNT_NAMES, [int(v) for v in out]))
        d.update(zchunk=zc.value, workgroups=wg.value)
        return d

    def compute_dt(self, level: int = -1) -> float:
        dt = self._C.c_double()
        self._check(self._lib.vh_blocks_compute_dt(self._h, self.cfl, int(level), self._C.byref(dt)))
        return dt.value

    def step(self, dt: float, level: int = -1):
        self._check(self._lib.vh_blocks_step(self._h, float(dt), int(level)))
        if level < 0:
            self.t += dt
            self.step_count += 1

    def evolve(self, tfinal: float = 0.0, max_steps: int = 0, subcycle: bool = False):
        C = self._C
        t = C.c_double(self.t)
        k = C.c_int64()
        self._check(self._lib.vh_blocks_run(self._h, self.cfl, float(tfinal), int(max_steps), int(bool(subcycle)),
                                            C.byref(t), C.byref(k)))
        self.t = t.value
        self.step_count += k.value
        return self.t, k.value

    def totals(self) -> np.ndarray:
        out = (self._C.c_double * 5)()
        self._check(self._lib.vh_blocks_totals(self._h, out))
        return np.array(list(out))

    def sync(self):
        self._check(self._lib.vh_blocks_sync(self._h))

    def indicator(self, cells: bool = False):
        """density_gradient_indicator (src/driver.jl:400-414) of every leaf, on the
        device. Returns the per-block maximum (nblocks,), and with cells=True also
        the per-leaf values (ncell,)."""
        C = self._C
        dp = C.POINTER(C.c_double)
        bmax = np.empty(self.mesh.nblocks)
        cell = np.empty(self.ncell) if cells else None
        self._check(self._lib.vh_blocks_indicator(self._h, bmax.ctypes.data_as(dp),
                                                  cell.ctypes.data_as(dp) if cells else None))
        return (bmax, cell) if cells else bmax

    def remap_from(self, src: "BlockSimulation", src_block):
        """Replace this simulation's state by `src`'s, transferred device to device
        onto this mesh: block b takes the cells of src block src_block[b] (the same
        block or its ancestor) by injection (prolong!, lib/RefMesh/src/RefMesh.jl:284-302).
        The clock is taken over from src."""
        C = self._C
        sb = np.ascontiguousarray(src_block, dtype=np.int32)
        if sb.shape != (self.mesh.nblocks,):
            raise ValueError(f"src_block must hold one source block per block ({self.mesh.nblocks})")
        self._check(self._lib.vh_blocks_remap(self._h, src._h, sb.ctypes.data_as(C.POINTER(C.c_int32))))
        self.t = src.t
        self.step_count = src.step_count

    def regrid(self, mesh: BlockMesh, src_block):
        """regrid! in place (src/driver.jl:440-462): this simulation moves onto
        `mesh` (same rank, maxlevel, h0 and gamma), block b taking the cells of
        current block src_block[b] (the same block or its ancestor) by injection
        (prolong!, lib/RefMesh/src/RefMesh.jl:284-302). The stream, precision,
        profile, launch shape and clock stay; device buffers are reused."""
        C = self._C
        if (mesh.rank, mesh.maxlevel, tuple(mesh.h0), mesh.gamma) != (self.mesh.rank, self.mesh.maxlevel,
                                                                     tuple(self.mesh.h0), self.mesh.gamma):
            raise ValueError("regrid keeps rank, maxlevel, h0 and gamma")
        sb = np.ascontiguousarray(src_block, dtype=np.int32)
        if sb.shape != (mesh.nblocks,):
            raise ValueError(f"src_block must hold one source block per block ({mesh.nblocks})")
        _, lv, dm, og, bc, keep = _mesh_args(mesh)
        self._check(self._lib.vh_blocks_regrid(self._h, mesh.nblocks, lv, dm, og, bc,
                                               sb.ctypes.data_as(C.POINTER(C.c_int32))))
        self.mesh = mesh
        self.ncell = int(self._lib.vh_blocks_ncell(self._h))

    def time_steps(self, nsteps: int, method: str = "graph", stage2: bool = True) -> dict:
        """Time nsteps single-rate steps: method 'persistent' (eager), 'per_block'
        (one launch per block per stage) or 'graph' (persistent, HIP graph)."""
        C = self._C
        code = {"persistent": 0, "per_block": 1, "graph": 2}[method]
        a, b = C.c_double(), C.c_double()
        self._check(self._lib.vh_blocks_time_steps(self._h, self.cfl, int(nsteps), code, C.byref(a),
                                                   C.byref(b) if stage2 else None))
        return {"step_ms": a.value, "stage2_ms": b.value if stage2 else None}
