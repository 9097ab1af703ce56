"""The host passes of the tier-1 seam (vespa.jl_amd/batched.py: gather of face
states in RefMesh's face order, scatter of ±F·area), checked on the CPU with the
oracle's per-face HLLC standing in for the GPU batch: gather → oracle fluxes →
scatter equals the oracle's literal face walk (accumulate_flux!) bit for bit on
every boundary mix, Float64 and Float32 (the reference's own gate:
test/test_amr_ka_flux.jl:74-95, batched ≡ native)."""
import numpy as np
import pytest

from conftest import load_vespa
from helpers import random_state

V = load_vespa()
from vespa_jl_amd import batched as B  # noqa: E402

NAMES = {0: "periodic", 1: "outflow", 2: "reflecting"}
CASES = [
    ((17,), [(1, 1)]),
    ((12,), [(2, 1)]),
    ((9,), [(0, 0)]),
    ((10, 7), [(1, 2), (0, 0)]),
    ((6, 5, 4), [(0, 0), (2, 1), (1, 2)]),
    ((5, 6, 3), [(0, 0), (0, 0), (0, 0)]),
]


def _areas(widths):
    out = []
    for a in range(len(widths)):
        ar = 1.0
        for d in range(len(widths)):
            if d != a:
                ar *= widths[d]
        out.append(ar)
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dims,bcs", CASES)
def test_gather_oracle_flux_scatter_equals_face_walk(O, dims, bcs, dtype):
    g = O.make_grid(dims, [(0.0, 1.0)] * len(dims), [(NAMES[a], NAMES[b]) for a, b in bcs], 1.4)
    U = random_state(tuple(reversed(dims)), seed=sum(dims), dtype=dtype)
    tab = B.FaceTable(dims, bcs)
    WL, WR, ax = B.face_states(tab, U, 1.4)
    F = O.face_flux(WL, WR, ax, 1.4)
    acc = B.scatter(tab, F, _areas([1.0 / n for n in dims]), dtype).reshape(U.shape)
    assert np.array_equal(acc, O.accumulate_flux(g, U))


def test_face_count_and_order():
    tab = B.FaceTable((4, 3), [(1, 1), (0, 0)])
    # x: 4 hi faces per row + 1 lo boundary face per row; y periodic: one face per cell
    assert tab.left.size == 3 * (4 + 1) + 12
    # first cell emits: x hi, x lo (boundary), y hi
    assert (tab.left[:3].tolist(), tab.right[:3].tolist(), tab.axis[:3].tolist()) == ([0, -1, 0], [1, 0, 4], [0, 0, 1])
