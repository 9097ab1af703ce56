"""The tier-1 seam: accumulate_flux! through a flux backend (src/ka_flux.jl:73-162).

When ``sim.flux`` holds a flux backend (``HIPFlux()``), a step runs the
reference's batched form instead of the fused device kernels:

  pass 1  gather the reconstructed left/right primitive states of every face in
          RefMesh's face order (for_each_face, lib/RefMesh/src/RefMesh.jl:187-197;
          _face_states, src/ka_flux.jl:75-91);
  pass 2  compute_face_fluxes!(backend, model, F, WL, WR, axisv) — the HLLC batch
          on the GPU through vh_face_flux_f64/_f32;
  pass 3  scatter ±F·area into both cells in the same face order
          (_scatter_face!, src/ka_flux.jl:94-126),

with the RK stages of step! (src/driver.jl:333-359) around it. Passes 1 and 3
run on the host, as the reference's do; the state crosses PCIe every stage, so
this path exists for the drop-in contract and its whole-run gate
(test/gpu/test_ka_flux_gpu.jl:93-102), not for speed — the fused device path
(``sim.flux = None``, the default) computes the same values without the host.
Arithmetic follows the reference operation by operation (numpy elementwise
arithmetic is IEEE and never fused), so both paths agree bit for bit.
"""
from __future__ import annotations

import numpy as np

from . import _lib

PER, OUT, REF = _lib.BC_PERIODIC, _lib.BC_OUTFLOW, _lib.BC_REFLECTING


def _cons2prim(U, g):
    """src/eos.jl:8-25 over (5, n)."""
    T = U.dtype.type
    r = U[0]
    inv = T(1) / r
    vx, vy, vz = U[1] * inv, U[2] * inv, U[3] * inv
    kin = (T(0.5) * r) * ((vx * vx + vy * vy) + vz * vz)
    p = (g - T(1)) * (U[4] - kin)
    p = np.maximum(p, T(1e-12) * kin)
    return np.stack([r, vx, vy, vz, p])


def _ghost(W, bc, axis):
    """src/driver.jl:117-122 ghost_state over (5, m): reflecting flips the normal velocity."""
    if bc == REF:
        W = W.copy()
        W[axis + 1] = -W[axis + 1]
    return W


def _minmod(a, b):
    """src/reconstruct.jl:6-12 (ties → b)."""
    T = a.dtype.type
    out = np.where(np.abs(a) < np.abs(b), a, b)
    return np.where(a * b <= 0, T(0), out)


class FaceTable:
    """RefMesh's face walk of a uniform mesh, vectorised: faces in emission order
    (per cell, column-major; per axis the hi face, then the lo face when it is a
    domain boundary). left/right are cell indices or -1 for a domain boundary."""

    def __init__(self, dims, bcs):
        dims = tuple(int(d) for d in dims)
        rank = len(dims)
        n = int(np.prod(dims))
        c = np.arange(n, dtype=np.int64)
        strides = [int(np.prod(dims[:d])) for d in range(rank)]
        keys, lefts, rights, axes = [], [], [], []
        self.nb = []   # per axis: (lo neighbour or -1, hi neighbour or -1) of every cell
        for d in range(rank):
            i = (c // strides[d]) % dims[d]
            per = bcs[d][0] == PER
            hi = np.where(i < dims[d] - 1, c + strides[d], (c - (dims[d] - 1) * strides[d]) if per else -1)
            lo = np.where(i > 0, c - strides[d], (c + (dims[d] - 1) * strides[d]) if per else -1)
            self.nb.append((lo, hi))
            # hi face of every cell
            keys.append((c * rank + d) * 2)
            lefts.append(c)
            rights.append(hi)
            axes.append(np.full(n, d))
            if not per:   # lo face at the domain boundary, right after the hi face
                sel = i == 0
                keys.append((c[sel] * rank + d) * 2 + 1)
                lefts.append(np.full(int(sel.sum()), -1))
                rights.append(c[sel])
                axes.append(np.full(int(sel.sum()), d))
        order = np.argsort(np.concatenate(keys), kind="stable")
        self.left = np.concatenate(lefts)[order]
        self.right = np.concatenate(rights)[order]
        self.axis = np.concatenate(axes)[order]
        self.dims, self.bcs, self.rank, self.ncell = dims, bcs, rank, n


def face_states(table: FaceTable, U: np.ndarray, gamma: float):
    """Pass 1: (WL, WR, axisv) as (nf, 5), (nf, 5), (nf,) int32 (axis 1-based)."""
    U = U.reshape(5, -1)
    T = U.dtype.type
    g = T(gamma)
    W = _cons2prim(U, g)
    hi_val, lo_val = [], []
    for d in range(table.rank):
        lo, hi = table.nb[d]
        Wl = np.where(lo >= 0, W[:, np.maximum(lo, 0)], _ghost(W, table.bcs[d][0], d))
        Wr = np.where(hi >= 0, W[:, np.maximum(hi, 0)], _ghost(W, table.bcs[d][1], d))
        s = _minmod(W - Wl, Wr - W)                         # _plm_slope (src/driver.jl:145-159)
        hi_val.append(W + T(0.5) * s)                       # _face_value :hi (src/driver.jl:227-233)
        lo_val.append(W + T(-0.5) * s)                      # _face_value :lo
    nf = table.left.size
    WL = np.empty((5, nf), dtype=U.dtype)
    WR = np.empty((5, nf), dtype=U.dtype)
    for d in range(table.rank):
        on = table.axis == d
        L, R = table.left[on], table.right[on]
        wl = np.where(L >= 0, hi_val[d][:, np.maximum(L, 0)], 0)
        wr = np.where(R >= 0, lo_val[d][:, np.maximum(R, 0)], 0)
        # boundary sides: the ghost of the interior face state (_boundary_ghost)
        bl, br = L < 0, R < 0
        wl[:, bl] = _ghost(wr[:, bl], table.bcs[d][0], d)
        wr[:, br] = _ghost(wl[:, br], table.bcs[d][1], d)
        WL[:, on] = wl
        WR[:, on] = wr
    return np.ascontiguousarray(WL.T), np.ascontiguousarray(WR.T), (table.axis + 1).astype(np.int32)


def scatter(table: FaceTable, F: np.ndarray, areas, dtype) -> np.ndarray:
    """Pass 3: acc (5, n) with acc[i] += F·a, acc[j] += F·(−a) in face order."""
    T = np.dtype(dtype).type
    aT = np.array([T(a) for a in areas], dtype=dtype)[table.axis]
    L, R = table.left, table.right
    # contributions in face order: left (+F·a) before right (F·(−a)) of the same face
    cells = np.stack([L, R], axis=1).ravel()
    keep = cells >= 0
    acc = np.zeros((5, table.ncell), dtype=dtype)
    for f in range(5):
        vals = np.stack([F[:, f] * aT, F[:, f] * (-aT)], axis=1).ravel()
        np.add.at(acc[f], cells[keep], vals[keep])
    return acc


def accumulate_flux_batched(table: FaceTable, U: np.ndarray, gamma: float, areas, backend, model) -> np.ndarray:
    """_accumulate_flux_batched! (src/ka_flux.jl:128-162) on a uniform mesh."""
    from .hydro import compute_face_fluxes
    WL, WR, axisv = face_states(table, U, gamma)
    F = np.empty_like(WL)
    compute_face_fluxes(backend, model, F, WL, WR, axisv)
    return scatter(table, F, areas, U.dtype)


def step_batched(sim, dt: float) -> None:
    """step!(sim, dt) (src/driver.jl:333-359) with accumulate_flux! through
    sim.flux: download, two batched stages, upload."""
    mesh = sim.mesh
    key = (tuple(mesh.dims), tuple(map(tuple, sim._bc_pairs)))
    if getattr(sim, "_face_table_key", None) != key:
        sim._face_table = FaceTable(mesh.dims, sim._bc_pairs)
        sim._face_table_key = key
    tab = sim._face_table
    T = sim.eltype.type
    w = mesh.cell_width()
    areas = []
    for a in range(mesh.rank):
        ar = 1.0
        for d in range(mesh.rank):
            if d != a:
                ar *= w[d]
        areas.append(ar)
    invV = T(dt) / (T(1.0) * T(mesh.cell_volume()))
    u0 = sim.state().reshape(5, -1)
    acc = accumulate_flux_batched(tab, u0, sim.problem.gamma, areas, sim.flux, sim.model)
    U1 = u0 - invV * acc                                       # _euler_apply!
    acc = accumulate_flux_batched(tab, U1, sim.problem.gamma, areas, sim.flux, sim.model)
    half = T(0.5)
    U2 = half * u0 + half * (U1 - invV * acc)                 # _rk2_combine!
    sim.upload(U2.reshape(sim.shape))
