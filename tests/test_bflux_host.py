"""Boundary-flux register layout and Enzo's face-plane raster (host logic, no
device): the dense register's plane/lookup against the oracle's slot layout,
and bflux_plane (src/reflux.jl:169-191) against a direct walk of the plane."""
import numpy as np

from conftest import load_vespa
from helpers import random_state

V = load_vespa()


class _FakeSim:
    def __init__(self, dims, eltype=np.float64):
        self.mesh = V.UniformMesh(tuple(dims), tuple((0.0, 1.0) for _ in dims))
        self.eltype = np.dtype(eltype)


def _register(dims, flux):
    r = V.BoundaryFluxRegister.__new__(V.BoundaryFluxRegister)
    r.dims = tuple(dims)
    r.flux = flux
    return r


def test_register_planes_follow_the_oracle_slot_layout(O):
    dims = (5, 4, 3)
    g = O.make_grid(dims, [(0.0, 1.0)] * 3, [("outflow", "reflecting"), ("reflecting", "outflow"),
                                              ("outflow", "outflow")], 1.4)
    U = random_state(tuple(reversed(dims)), seed=3, vmax=0.4)
    nb = O.bface_count(g)
    bf = np.zeros((nb, 5))
    O.step(g, U, 1e-3, bflux=bf)
    reg = _register(dims, bf)
    # the z-lo face of cell (i, j, 1) (1-based) is the (x, y) plane entry at [j-1, i-1]
    for axis, side in ((1, "lo"), (1, "hi"), (2, "lo"), (3, "hi")):
        P = reg.plane(axis, side)
        assert P.shape[-1] == 5 and P.size == 5 * (np.prod(dims) // dims[axis - 1])
    # every recorded face carries a nonzero mass flux only on the faces the flow crosses;
    # check lookup against the slot formula: planes (x lo, x hi, y lo, y hi, z lo, z hi)
    n = int(np.prod(dims))
    for (axis, side, cell) in (((1, "hi", (5, 2, 3))), (2, "lo", (4, 1, 2)), (3, "hi", (1, 4, 3))):
        base = sum(2 * (n // dims[e]) for e in range(axis - 1)) + (n // dims[axis - 1] if side == "hi" else 0)
        other = [e for e in range(3) if e != axis - 1]
        idx = (cell[other[0]] - 1) + (cell[other[1]] - 1) * dims[other[0]]
        np.testing.assert_array_equal(reg.lookup(axis, side, cell), bf[base + idx])


def test_bflux_plane_walks_the_face_column_major():
    rng = np.random.default_rng(0)
    store = {}

    def lookup(I):
        return store.get(I)
    for I in [(3, j, k) for j in range(1, 5) for k in range(1, 4)]:
        store[I] = rng.normal(size=5)
    # x face (dim 0) of a 3D grid: start/stop global indices, g0 the grid's first active cell
    start, stop, g0 = (10, 7, 2), (10, 10, 4), (8, 7, 2)
    out = V.bflux_plane(3, 0, start, stop, g0, 3, 2, 0.5, lookup)
    assert out.shape == (1 * 4 * 3,)
    lin = 0
    for k in range(3):                 # dims beyond the first vary slowest
        for j in range(4):
            I = (3, j + 1, k + 1)      # flux dim: flux_off; others: g - g0 + 1
            assert out[lin] == store[I][1] / 0.5
            lin += 1
    # comp 0 and missing entries give zeros
    assert not V.bflux_plane(3, 0, start, stop, g0, 3, 0, 0.5, lookup).any()
    assert not V.bflux_plane(3, 0, start, stop, g0, 9, 1, 0.5, lookup).any()
