"""Independent numpy restatement of one RefMesh SSP-RK2 step (test-only checker).

Written separately from the C oracle to cross-check it on small grids: faces are
computed per axis with vectorised numpy (ghost states synthesised per boundary
rule, src/driver.jl:117-159), and each cell then gathers its six face fluxes in
the order the reference's per-face scatter produces (derived from
lib/RefMesh/src/RefMesh.jl:187-197):

  1. lo faces whose partner cell is visited earlier (non-wrapped interior), z, y, x;
  2. per axis x, y, z: the cell's hi face, then its lo face if it is a domain
     boundary;
  3. lo faces wrapped by a periodic boundary, x, y, z.

The C oracle instead walks faces literally, so agreement pins that ordering.
numpy elementwise float arithmetic is IEEE and never fused, matching Julia.
"""
from __future__ import annotations

import numpy as np

PER, OUT, REF = 0, 1, 2


def cons2prim(U, g):
    T = U.dtype.type
    r = U[0]
    inv = T(1) / r
    vx, vy, vz = U[1] * inv, U[2] * inv, U[3] * inv
    kin = (T(0.5) * r) * ((vx * vx + vy * vy) + vz * vz)
    p = (g - T(1)) * (U[4] - kin)
    p = np.maximum(p, T(1e-12) * kin)
    return np.stack([r, vx, vy, vz, p])


def minmod(a, b):
    T = a.dtype.type
    out = np.where(np.abs(a) < np.abs(b), a, b)
    return np.where(a * b <= 0, T(0), out)


def energy(W, g):
    T = W.dtype.type
    return W[4] / (g - T(1)) + (T(0.5) * W[0]) * ((W[1] * W[1] + W[2] * W[2]) + W[3] * W[3])


def eflux(W, g, dim):
    T = W.dtype.type
    r, p = W[0], W[4]
    un = W[dim]
    E = energy(W, g)
    z = np.zeros_like(r)
    return np.stack([r * un,
                     (r * W[1]) * un + (p if dim == 1 else z),
                     (r * W[2]) * un + (p if dim == 2 else z),
                     (r * W[3]) * un + (p if dim == 3 else z),
                     (E + p) * un])


def hllc(WL, WR, g, dim):
    """Toro's HLLC with Davis speeds, evaluated for both sides then selected."""
    with np.errstate(all="ignore"):
        rL, rR, pL, pR = WL[0], WR[0], WL[4], WR[4]
        uL, uR = WL[dim], WR[dim]
        cL = np.sqrt(g * pL / rL)
        cR = np.sqrt(g * pR / rR)
        SL = np.minimum(uL - cL, uR - cR)
        SR = np.maximum(uL + cL, uR + cR)
        FL, FR = eflux(WL, g, dim), eflux(WR, g, dim)
        Ss = (pR - pL + rL * uL * (SL - uL) - rR * uR * (SR - uR)) / (rL * (SL - uL) - rR * (SR - uR))

        def star(W, F, SK):
            r, un, p = W[0], W[dim], W[4]
            E = energy(W, g)
            U = np.stack([r, r * W[1], r * W[2], r * W[3], E])
            fac = r * (SK - un) / (SK - Ss)
            vs = [W[1], W[2], W[3]]
            vs[dim - 1] = Ss
            Es = E / r + (Ss - un) * (Ss + p / (r * (SK - un)))
            Us = np.stack([fac, fac * vs[0], fac * vs[1], fac * vs[2], fac * Es])
            return F + SK * (Us - U)

        sL, sR = star(WL, FL, SL), star(WR, FR, SR)
        out = np.where(Ss >= 0, sL, sR)
        out = np.where(SR <= 0, FR, out)
        out = np.where(SL >= 0, FL, out)
    return out


def _ghost(W, bc, axis):
    G = W.copy()
    if bc == REF:
        G[axis + 1] = -G[axis + 1]
    return G


def _axis_faces(W, g, ax, dim, n_ax, bclo, bchi):
    """Face fluxes along array axis `ax` for the 1-based spatial axis `dim`.

    Returns F of shape with n_ax + 1 faces along that axis: face i sits between
    cells i-1 and i. For periodic axes face 0 and face n are the same wrap face."""
    spatial = dim - 1
    T = W.dtype.type
    take = lambda i: np.take(W, i, axis=ax)
    per = bclo == PER
    if per:
        Wext = np.concatenate([take([n_ax - 1]), W, take([0])], axis=ax)
    else:
        glo = _ghost(take([0]), bclo, spatial)
        ghi = _ghost(take([n_ax - 1]), bchi, spatial)
        Wext = np.concatenate([glo, W, ghi], axis=ax)
    # slopes of the real cells
    sl = lambda a, b: np.take(Wext, np.arange(a, b), axis=ax)
    s = minmod(sl(1, n_ax + 1) - sl(0, n_ax), sl(2, n_ax + 2) - sl(1, n_ax + 1))
    hi_val = W + T(0.5) * s
    lo_val = W + T(-0.5) * s
    # interior faces i = 1..n-1: L = hi_val[i-1], R = lo_val[i]
    L = np.take(hi_val, np.arange(0, n_ax - 1), axis=ax)
    R = np.take(lo_val, np.arange(1, n_ax), axis=ax)
    Fint = hllc(L, R, g, dim)
    if per:
        Lw = np.take(hi_val, [n_ax - 1], axis=ax)
        Rw = np.take(lo_val, [0], axis=ax)
        Fw = hllc(Lw, Rw, g, dim)
        F = np.concatenate([Fw, Fint, Fw], axis=ax)
    else:
        Rb = np.take(lo_val, [0], axis=ax)
        Lb = _ghost(Rb, bclo, spatial)
        F0 = hllc(Lb, Rb, g, dim)
        Lh = np.take(hi_val, [n_ax - 1], axis=ax)
        Rh = _ghost(Lh, bchi, spatial)
        Fn = hllc(Lh, Rh, g, dim)
        F = np.concatenate([F0, Fint, Fn], axis=ax)
    return F


def accumulate(U, dims, bcs, gamma, widths):
    """acc for state U of shape (5, nz, ny, nx) (rank-trimmed)."""
    rank = len(dims)
    T = U.dtype.type
    g = T(gamma)
    W = cons2prim(U, g)
    areas = []
    for a in range(rank):
        ar = 1.0
        for d in range(rank):
            if d != a:
                ar *= widths[d]
        areas.append(T(ar))
    faces = []
    for d in range(rank):
        ax = rank - d            # spatial axis d ↔ array axis
        F = _axis_faces(W, g, ax, d + 1, dims[d], bcs[d][0], bcs[d][1])
        lo = np.take(F, np.arange(0, dims[d]), axis=ax)
        hi = np.take(F, np.arange(1, dims[d] + 1), axis=ax)
        # lo-face kind per cell: 0 interior, 1 boundary, 2 wrapped
        idx = np.arange(dims[d])
        kind = np.where(idx > 0, 0, 2 if bcs[d][0] == PER else 1)
        shape = [1] * (rank + 1)
        shape[ax] = dims[d]
        faces.append((lo, hi, kind.reshape(shape)[0], areas[d]))
    acc = np.zeros_like(U)
    # 1. preceding interior lo faces, highest axis first
    for d in reversed(range(rank)):
        lo, hi, kind, a = faces[d]
        acc = np.where(kind == 0, acc + (-a) * lo, acc)
    # 2. hi face then boundary lo face, per axis
    for d in range(rank):
        lo, hi, kind, a = faces[d]
        acc = acc + a * hi
        acc = np.where(kind == 1, acc + (-a) * lo, acc)
    # 3. wrapped lo faces
    for d in range(rank):
        lo, hi, kind, a = faces[d]
        acc = np.where(kind == 2, acc + (-a) * lo, acc)
    return acc


def step(U, dims, bcs, gamma, widths, dt):
    """One SSP-RK2 step (src/driver.jl:333-359) returning the new state."""
    T = U.dtype.type
    V = widths[0]
    for w in widths[1:]:
        V = V * w
    invV = T(dt) / (T(1.0) * T(V))
    u0 = U.copy()
    U1 = u0 - invV * accumulate(u0, dims, bcs, gamma, widths)
    half = T(0.5)
    return half * u0 + half * (U1 - invV * accumulate(U1, dims, bcs, gamma, widths))
