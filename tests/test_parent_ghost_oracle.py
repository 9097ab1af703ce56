"""Pins of the oracle's ParentGhost boundary rule (lib/MeshInterface/src/MeshInterface.jl:86-99,
src/driver.jl:127-134, 145-159, 255-281): at a ParentGhost side the boundary ghost
is the closure's primitive state for (axis, side, cell), used both as the outer
neighbour of the boundary cell's slope and as the outer state of the boundary
face; the oracle takes those states from a per-face array in the boundary-flux
slot order.

* Against a larger domain: a flow supersonic along the axis, entering through
  the ParentGhost side from a uniform upstream region, on a grid that is the
  downstream part of a longer grid. Upstream of the face every cell holds the
  same state, keeps it (every flux into it is the upwind flux of that state)
  and reconstructs it with a zero slope, so the longer grid's outer neighbour
  and face state are exactly the ParentGhost state: one step of both agrees on
  the shared cells to the round-off of the boundary cells' summation order
  (bit for bit in 1D), for each axis and side, in 1D, 2D and 3D; a different
  ghost state moves the result by orders of magnitude more.
* Plane indexing: with a state uniform across the ParentGhost plane, moving the
  ghost states one cell along each transverse axis moves the result the same
  way (periodic transverse axes; equal up to the order of the wrapped faces).
"""
import numpy as np
import pytest

from conftest import load_vespa
from helpers import prim2cons

V = load_vespa()
G14 = 1.4


def _prim_field(dims, axis, sign, seed, gamma=G14):
    """Supersonic bulk flow along `axis` (speed 3c, direction `sign`) with smooth
    density/pressure and small transverse velocities."""
    rng = np.random.default_rng(seed)
    rank = len(dims)
    shape = tuple(dims[::-1])
    grids = np.meshgrid(*[np.arange(n) / n for n in dims[::-1]], indexing="ij")[::-1]
    ph = rng.uniform(0, 2 * np.pi, 8)
    s = sum(np.sin(2 * np.pi * grids[d] + ph[d]) for d in range(rank))
    rho = 1.0 + 0.2 * s / rank
    p = 1.0 + 0.1 * np.cos(ph[4] + s)
    c = np.sqrt(gamma * p / rho)
    v = [0.05 * np.sin(2 * np.pi * grids[d % rank] + ph[5 + d]) * np.ones(shape) for d in range(3)]
    v[axis] = sign * 3.0 * c
    return np.stack([rho, v[0], v[1], v[2], p]).reshape(5, -1)


@pytest.mark.parametrize("rank", [1, 2, 3])
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("side", [0, 1])
def test_parent_ghost_is_the_upstream_of_a_longer_domain(O, rank, axis, side):
    if axis >= rank:
        pytest.skip("axis beyond rank")
    base = [9, 6, 5][:rank]
    n, ext = base[axis], 4
    sign = 1.0 if side == 0 else -1.0          # flow enters through `side`
    dims_s = list(base)
    dims_b = list(base)
    dims_b[axis] = n + ext
    # γ = 2 and an upstream state (ρ 2, p 1, c 1, |u| 4, no transverse velocity)
    # whose fluxes are small integers, so the upstream cells' flux sums cancel
    # exactly in any order and the region stays uniform to the last bit
    gam = 2.0
    W_s = _prim_field(dims_s, axis, sign, seed=rank * 7 + axis * 2 + side, gamma=gam)
    Wu = np.array([2.0, 0.0, 0.0, 0.0, 1.0])
    Wu[1 + axis] = sign * 4.0
    Wb = np.empty((5,) + tuple(dims_b[::-1]))
    sl_s = [slice(None)] * rank
    sl_u = [slice(None)] * rank
    ax = rank - 1 - axis                        # numpy axis of `axis` in (.., z, y, x) order
    if side == 0:
        sl_s[ax], sl_u[ax] = slice(ext, None), slice(0, ext)
    else:
        sl_s[ax], sl_u[ax] = slice(0, n), slice(n, None)
    Wb[(slice(None),) + tuple(sl_s)] = W_s.reshape((5,) + tuple(dims_s[::-1]))
    Wb[(slice(None),) + tuple(sl_u)] = Wu.reshape((5,) + (1,) * rank)
    bcs_s, bcs_b = [], []
    for d in range(rank):
        if d == axis:
            pair = ["outflow", "outflow"]
            pair[side] = "parent"
            bcs_s.append(tuple(pair))
            bcs_b.append(("outflow", "outflow"))
        else:
            bcs_s.append(("periodic", "periodic"))
            bcs_b.append(("periodic", "periodic"))
    # unit cells on both grids (equal widths, areas and volumes)
    gs = O.make_grid(tuple(dims_s), [(0.0, float(n_)) for n_ in dims_s], bcs_s, gam)
    gb = O.make_grid(tuple(dims_b), [(0.0, float(n_)) for n_ in dims_b], bcs_b, gam)
    nbf = sum(2 * int(np.prod(dims_s)) // dims_s[d] for d in range(rank))
    pg = np.random.default_rng(1).uniform(0.5, 1.5, (nbf, 5))   # unused slots hold noise
    for (d, s_, first, shape) in O.bface_planes(gs):
        if d == axis and s_ == side:
            pg[first:first + shape[0] * shape[1]] = Wu
    O.set_parent_ghost(gs, pg)
    Us = np.ascontiguousarray(prim2cons(W_s, gam))
    Ub = np.ascontiguousarray(prim2cons(Wb.reshape(5, -1), gam))
    dt = 0.5 * min(O.compute_dt(gs, Us, 0.4), O.compute_dt(gb, Ub, 0.4))
    O.step(gs, Us, dt)
    O.step(gb, Ub, dt)
    Ubv = Ub.reshape((5,) + tuple(dims_b[::-1]))[(slice(None),) + tuple(sl_s)].reshape(5, -1)
    # the boundary cells add their boundary face after their hi face (RefMesh's
    # walk), where the longer grid adds that face first: round-off only
    scale = np.abs(Ubv).max(axis=1, keepdims=True)
    assert (np.abs(Us - Ubv) <= 1e-14 * scale).all()
    # the upstream region stayed uniform (the premise of the comparison)
    up = Ub.reshape((5,) + tuple(dims_b[::-1]))[(slice(None),) + tuple(sl_u)].reshape(5, -1)
    assert np.array_equal(up, np.repeat(prim2cons(Wu[:, None], gam), up.shape[1], axis=1))
    # the ghost states matter: another upstream state changes the result
    pg2 = pg.copy()
    for (d, s_, first, shape) in O.bface_planes(gs):
        if d == axis and s_ == side:
            pg2[first:first + shape[0] * shape[1], 0] *= 1.1
    O.set_parent_ghost(gs, pg2)
    U2 = np.ascontiguousarray(prim2cons(W_s, gam))
    O.step(gs, U2, dt)
    assert np.abs(U2 - Us).max() > 1e-4 * np.abs(Us).max()


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_parent_ghost_plane_indexing_moves_with_the_states(O, axis):
    dims = [8, 6, 5]
    n = dims[axis]
    others = [d for d in range(3) if d != axis]
    # a state that varies along `axis` only
    x = (np.arange(n) + 0.5) / n
    rho_1d = 1.0 + 0.3 * np.sin(2 * np.pi * x)
    shape = tuple(dims[::-1])
    idx = [np.newaxis] * 3
    idx[2 - axis] = slice(None)
    rho = np.broadcast_to(rho_1d[tuple(idx)], shape)
    W = np.stack([rho, 0.1 * np.ones(shape), -0.05 * np.ones(shape), 0.02 * np.ones(shape), np.ones(shape)])
    U0 = np.ascontiguousarray(prim2cons(W.reshape(5, -1), G14))
    bcs = []
    for d in range(3):
        bcs.append(("parent", "outflow") if d == axis else ("periodic", "periodic"))
    g = O.make_grid(tuple(dims), [(0.0, 1.0)] * 3, bcs, G14)
    planes = {(d, s): (first, shp) for (d, s, first, shp) in O.bface_planes(g)}
    first, shp = planes[(axis, 0)]
    nbf = sum(2 * int(np.prod(dims)) // dims[d] for d in range(3))
    rng = np.random.default_rng(4)
    plane = np.stack([rng.uniform(0.8, 1.2, shp[::-1]), rng.normal(0, 0.1, shp[::-1]),
                      rng.normal(0, 0.1, shp[::-1]), rng.normal(0, 0.1, shp[::-1]),
                      rng.uniform(0.8, 1.2, shp[::-1])], axis=-1)        # (slow, fast, 5)

    def run(pl):
        pg = np.zeros((nbf, 5))
        pg[first:first + shp[0] * shp[1]] = pl.reshape(-1, 5)
        O.set_parent_ghost(g, pg)
        U = U0.copy()
        O.step(g, U, 0.2 / max(dims))
        return U.reshape((5,) + shape)

    R = run(plane)
    for k, e in enumerate(others):             # plane axes: (slow = others[1], fast = others[0])
        moved = np.roll(plane, 1, axis=1 - k)
        Rm = run(moved)
        expect = np.roll(R, 1, axis=1 + (2 - e))
        assert np.allclose(Rm, expect, rtol=1e-12, atol=1e-14)
        assert not np.allclose(Rm, R, rtol=1e-6)


def test_parent_ghost_needs_its_states(O):
    g = O.make_grid((8, 6), [(0.0, 1.0)] * 2, [("parent", "outflow"), ("periodic", "periodic")], G14)
    U = np.ascontiguousarray(prim2cons(np.tile(np.array([[1.0], [0.1], [0.0], [0.0], [1.0]]), 48), G14))
    with pytest.raises(ValueError):
        O.step(g, U, 0.01)
