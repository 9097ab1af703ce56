"""Pins of the CPU oracle against what the mathematics and the reference's own
tests fix (no golden vectors exist for this path in the reference).

Each check is chosen so that a plausible restatement mistake (a dropped term, a
wrong sign or index, a transposed operand, a changed accumulation order) fails it.
"""
import numpy as np
import pytest

from helpers import prim2cons, random_faces, random_state, smooth_state
import np_restatement as NR

G14 = 1.4

# --------------------------------------------------------------------------
# exact Riemann solver (src/exact_riemann.jl) — Toro, Riemann Solvers and
# Numerical Methods for Fluid Dynamics, table 4.3 (γ = 1.4): star states of the
# five standard tests.
TORO = [
    ((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 0.30313, 0.92745, 0.42632, 0.26557),
    ((1.0, -2.0, 0.4), (1.0, 2.0, 0.4), 0.00189, 0.00000, 0.02185, 0.02185),
    ((1.0, 0.0, 1000.0), (1.0, 0.0, 0.01), 460.894, 19.5975, 0.57506, 5.99924),
    ((1.0, 0.0, 0.01), (1.0, 0.0, 100.0), 46.0950, -6.19633, 5.99242, 0.57511),
    ((5.99924, 19.5975, 460.894), (5.99242, -6.19633, 46.0950), 1691.64, 8.68975, 14.2823, 31.0426),
]


@pytest.mark.parametrize("case", range(5))
def test_exact_riemann_toro_star_states(O, case):
    WL, WR, ps, us, rsl, rsr = TORO[case]
    p, u = O.star_state(WL, WR, G14)
    assert p == pytest.approx(ps, rel=2e-4, abs=2e-5)
    assert u == pytest.approx(us, rel=2e-4, abs=2e-5)
    # densities just either side of the contact (sample at u* ∓ tiny)
    rl = O.exact_riemann_sample(WL, WR, G14, u - 1e-9)[0]
    rr = O.exact_riemann_sample(WL, WR, G14, u + 1e-9)[0]
    assert rl == pytest.approx(rsl, rel=2e-4, abs=2e-5)
    assert rr == pytest.approx(rsr, rel=2e-4, abs=2e-5)


def test_exact_riemann_far_field_and_fan(O):
    WL, WR = (1.0, 0.0, 1.0), (0.125, 0.0, 0.1)
    assert tuple(O.exact_riemann_sample(WL, WR, G14, -10.0)) == WL
    assert tuple(O.exact_riemann_sample(WL, WR, G14, 10.0)) == WR
    # inside the left rarefaction the Riemann invariant u + 2c/(γ-1) is constant
    cL = np.sqrt(G14)
    for s in (-1.0, -0.5, -0.2):
        r, u, p = O.exact_riemann_sample(WL, WR, G14, s)
        c = np.sqrt(G14 * p / r)
        assert u + 2 * c / (G14 - 1) == pytest.approx(0.0 + 2 * cL / (G14 - 1), rel=1e-12)
        assert p / r ** G14 == pytest.approx(1.0, rel=1e-12)      # isentropic fan


# --------------------------------------------------------------------------
# EOS (src/eos.jl) and minmod (src/reconstruct.jl)

def test_eos_round_trip_and_floor(O):
    rng = np.random.default_rng(3)
    for _ in range(200):
        W = np.array([rng.uniform(0.1, 5), *rng.uniform(-3, 3, 3), rng.uniform(0.01, 10)])
        U = O.prim2cons(W, G14)
        assert U[4] == W[4] / (G14 - 1) + (0.5 * W[0]) * ((W[1] ** 2 + W[2] ** 2) + W[3] ** 2)
        W2 = O.cons2prim(U, G14)
        np.testing.assert_allclose(W2, W, rtol=1e-12, atol=1e-13)
    # E below the kinetic energy: the floor max(p, 1e-12·kinetic) engages
    U = np.array([1.0, 2.0, 0.0, 0.0, 1.0])            # kinetic = 2 > E
    W = O.cons2prim(U, G14)
    assert W[4] == 1e-12 * 2.0
    # float32 form: pressure evaluated in double, then rounded
    U32 = np.array([1.0, 1e-3, 0.0, 0.0, 5e-7 + 1e-7], dtype=np.float32)
    Wd = O.cons2prim(U32, G14, dtype=np.float32, floor64=True)
    kin = np.float32(np.float32(0.5) * np.float32(1.0)) * (np.float32(1e-3) * np.float32(1e-3))
    assert Wd[4] == np.float32((G14 - 1.0) * (float(U32[4]) - float(kin)))


@pytest.mark.parametrize("a,b,want", [
    (1.0, 2.0, 1.0), (2.0, 1.0, 1.0), (-1.0, -2.0, -1.0), (-2.0, -1.0, -1.0),
    (1.0, -1.0, 0.0), (-1.0, 1.0, 0.0), (0.0, 5.0, 0.0), (5.0, 0.0, 0.0),
    (2.0, 2.0, 2.0), (-3.0, -3.0, -3.0)])
def test_minmod_table(O, a, b, want):
    assert O.minmod(a, b) == want


def test_minmod_tie_and_underflow(O):
    # equal values return b (= a); opposite signs of equal magnitude give 0
    assert O.minmod(0.5, 0.5) == 0.5
    assert O.minmod(0.5, -0.5) == 0.0
    # the reference tests the product a·b ≤ 0: when it underflows to 0 the slope
    # is zeroed even though both differences are positive
    assert O.minmod(1e-300, 1e-300) == 0.0
    assert O.minmod(1e-150, 1e-150) == 1e-150


# --------------------------------------------------------------------------
# HLLC special cases (src/riemann.jl:33-88)

def _euler_flux_textbook(W, g, dim):
    r, v, p = W[0], np.array(W[1:4]), W[4]
    E = p / (g - 1) + 0.5 * r * (v @ v)
    un = v[dim - 1]
    F = np.array([r * un, *(r * v * un), (E + p) * un])
    F[dim] += p
    return F


def test_hllc_supersonic_is_upwind_flux_bitwise(O):
    W = np.array([1.0, 5.0, 0.3, -0.2, 1.0])
    for dim in (1, 2, 3):
        Wd = W.copy()
        Wd[1:4] = [0.3, -0.2, 0.1]
        Wd[dim] = 5.0
        F = O.hllc_flux(Wd, Wd * [1.1, 1, 1, 1, 0.9], G14, dim)
        assert np.array_equal(F, O.euler_flux(Wd, G14, dim))           # SL ≥ 0 → FL
        Wn = Wd.copy()
        Wn[dim] = -5.0
        F = O.hllc_flux(Wn * [1.1, 1, 1, 1, 0.9], Wn, G14, dim)
        assert np.array_equal(F, O.euler_flux(Wn, G14, dim))           # SR ≤ 0 → FR


def test_hllc_consistency_and_physical_flux(O):
    rng = np.random.default_rng(5)
    for _ in range(100):
        W = np.array([rng.uniform(0.2, 3), *rng.uniform(-1, 1, 3), rng.uniform(0.1, 3)])
        for dim in (1, 2, 3):
            F = O.hllc_flux(W, W, G14, dim)
            np.testing.assert_allclose(F, _euler_flux_textbook(W, G14, dim), rtol=1e-13, atol=1e-13)


def test_hllc_stationary_contact(O):
    WL = np.array([1.0, 0.0, 0.0, 0.0, 1.0])
    WR = np.array([0.125, 0.0, 0.0, 0.0, 1.0])
    for dim in (1, 2, 3):
        F = O.hllc_flux(WL, WR, G14, dim)
        want = np.zeros(5)
        want[dim] = 1.0
        np.testing.assert_allclose(F, want, rtol=0, atol=1e-15)


def test_hllc_matches_independent_numpy_restatement_bitwise(O):
    WL, WR, ax = random_faces(4096, seed=11)
    F = O.face_flux(WL, WR, ax, G14)
    for dim in (1, 2, 3):
        m = ax == dim
        Fn = NR.hllc(WL[m].T.copy(), WR[m].T.copy(), np.float64(G14), dim).T
        assert np.array_equal(F[m], Fn)


def test_hllc_mirror_symmetry(O):
    """Reflecting both states through the face negates the mass and energy flux
    and keeps the normal momentum flux (to round-off)."""
    rng = np.random.default_rng(7)
    for _ in range(200):
        WL = np.array([rng.uniform(0.2, 3), *rng.uniform(-2, 2, 3), rng.uniform(0.1, 3)])
        WR = np.array([rng.uniform(0.2, 3), *rng.uniform(-2, 2, 3), rng.uniform(0.1, 3)])
        dim = 1
        mL, mR = WR.copy(), WL.copy()
        mL[1] = -mL[1]
        mR[1] = -mR[1]
        F = O.hllc_flux(WL, WR, G14, dim)
        Fm = O.hllc_flux(mL, mR, G14, dim)
        np.testing.assert_allclose(Fm, [-F[0], F[1], -F[2], -F[3], -F[4]], rtol=1e-12, atol=1e-12)


# --------------------------------------------------------------------------
# Whole-path invariants and the reference's own thresholds

def _sod(O, n, dtype=np.float64):
    g = O.make_grid((n,), ((0.0, 1.0),), "outflow", G14)
    x = 0.0 + (np.arange(1, n + 1) - 0.5) * (1.0 / n)
    W = np.zeros((5, n))
    W[0] = np.where(x < 0.5, 1.0, 0.125)
    W[4] = np.where(x < 0.5, 1.0, 0.1)
    U = prim2cons(W.astype(dtype), G14, dtype)
    return g, x, U


def _sod_l1(O, x, U, t):
    rho = U[0].astype(np.float64)
    ex = np.array([O.exact_riemann_sample((1, 0, 1), (0.125, 0, 0.1), G14, (xi - 0.5) / t)[0] for xi in x])
    return np.abs(rho - ex).mean()


def test_sod_256_matches_reference_thresholds(O):
    """test/test_sod.jl:20-44: L1(ρ) < 0.01, mass/energy rtol 1e-12, bounds."""
    g, x, U = _sod(O, 256)
    t0 = O.totals(g, U)
    t, k = O.evolve(g, U, 0.2, 0.4)
    t1 = O.totals(g, U)
    assert t == pytest.approx(0.2, abs=1e-15)
    assert t1[0] == pytest.approx(t0[0], rel=1e-12)
    assert t1[4] == pytest.approx(t0[4], rel=1e-12)
    l1 = _sod_l1(O, x, U, t)
    assert l1 < 0.01
    W = np.array([O.cons2prim(U[:, i], G14) for i in range(256)]).T
    assert (W[0] > 0).all() and (W[4] > 0).all()
    assert W[0].max() <= 1.0 + 1e-6 and W[0].min() >= 0.125 - 1e-6


def test_sod_float32_precision_genericity(O):
    """test/test_sod.jl:46-70: L1₃₂ < 0.02 and < 5·L1₆₄ (decoupled f64 geometry)."""
    g, x, U64 = _sod(O, 256)
    t64, _ = O.evolve(g, U64, 0.2, 0.4)
    l1_64 = _sod_l1(O, x, U64, t64)
    for floor64 in (False, True):
        _, _, U32 = _sod(O, 256, np.float32)
        t32, _ = O.evolve(g, U32, 0.2, 0.4, floor64=floor64)
        assert np.isfinite(U32).all() and (U32[0] > 0).all()
        l1_32 = _sod_l1(O, x, U32, t32)
        assert l1_32 < 0.02 and l1_32 < 5 * l1_64


def test_sod_mirror_symmetry(O):
    """The tube reversed end to end gives the mirrored solution (to round-off)."""
    g, x, U = _sod(O, 200)
    Um = U[:, ::-1].copy()
    Um[1] = -Um[1]
    dts = O.run_steps(g, U, 40, 0.4)
    dtm = O.run_steps(g, Um, 40, 0.4)
    np.testing.assert_allclose(dts, dtm, rtol=1e-13)
    np.testing.assert_allclose(Um[0, ::-1], U[0], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(-Um[1, ::-1], U[1], rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("dims,bcs", [
    ((16, 12, 10), "periodic"), ((24, 20), "periodic"), ((13, 9, 7), "reflecting"),
    ((12, 10), [("reflecting", "reflecting"), ("periodic", "periodic")])])
def test_conservation_closed_boxes(O, dims, bcs):
    g = O.make_grid(dims, [(0.0, 1.0)] * len(dims), bcs, 5.0 / 3.0)
    U = smooth_state(dims, seed=2)
    t0 = O.totals(g, U)
    O.run_steps(g, U, 5, 0.3)
    t1 = O.totals(g, U)
    np.testing.assert_allclose(t1[[0, 4]], t0[[0, 4]], rtol=1e-13)
    if bcs == "periodic":
        np.testing.assert_allclose(t1[1:4], t0[1:4], rtol=0, atol=1e-14)


def test_uniform_state_is_preserved(O):
    dims = (8, 6, 5)
    g = O.make_grid(dims, [(0.0, 1.0)] * 3, "periodic", G14)
    U = prim2cons(np.broadcast_to(np.array([1.3, 0.2, -0.1, 0.05, 0.7])[:, None, None, None],
                                  (5, 5, 6, 8)).copy(), G14)
    U0 = U.copy()
    O.run_steps(g, U, 3, 0.4)
    np.testing.assert_allclose(U, U0, rtol=1e-13, atol=1e-15)


def test_boundary_flux_identity(O):
    """test/test_boundary_flux.jl:9-36: Δmass ≡ Σ_lo bflux[ρ] − Σ_hi bflux[ρ]."""
    n = 64
    g = O.make_grid((n,), ((0.0, 1.0),), "outflow", G14)
    x = (np.arange(1, n + 1) - 0.5) / n
    W = np.stack([1.0 + x, np.ones(n), np.zeros(n), np.zeros(n), 2.0 * np.ones(n)])
    U = prim2cons(W, G14)
    m0 = U[0].sum() / n
    lo = hi = 0.0
    for _ in range(20):
        bf = np.zeros((O.bface_count(g), 5))
        O.step(g, U, 5e-4, bflux=bf)
        lo += bf[0, 0]
        hi += bf[1, 0]
    dm = U[0].sum() / n - m0
    assert abs(dm) > 1e-3
    assert dm == pytest.approx(lo - hi, abs=1e-12)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dims,bcs", [
    ((7,), [(1, 1)]), ((8,), [(0, 0)]), ((9,), [(2, 1)]), ((6, 5), [(0, 0), (2, 1)]),
    ((5, 4, 3), [(1, 2), (0, 0), (2, 2)]), ((4, 4, 4), [(0, 0)] * 3), ((3, 5, 4), [(2, 1), (1, 2), (0, 0)])])
def test_oracle_step_equals_numpy_restatement_bitwise(O, dims, bcs, dtype):
    """Literal face walk (oracle) ≡ per-cell gather in the derived order (numpy)."""
    names = {0: "periodic", 1: "outflow", 2: "reflecting"}
    U = random_state(tuple(reversed(dims)), seed=sum(dims), dtype=dtype)
    g = O.make_grid(dims, [(0.0, 1.0)] * len(dims), [(names[a], names[b]) for a, b in bcs], G14)
    widths = [1.0 / d for d in dims]
    acc_o = O.accumulate_flux(g, U)
    acc_n = NR.accumulate(U, dims, bcs, G14, widths)
    assert np.array_equal(acc_o, acc_n)
    Uo = U.copy()
    O.step(g, Uo, 0.01)
    assert np.array_equal(Uo, NR.step(U, dims, bcs, G14, widths, 0.01))


def test_compute_dt_definition(O):
    dims = (6, 5, 4)
    g = O.make_grid(dims, [(0.0, 1.0), (0.0, 2.0), (0.0, 0.5)], "outflow", G14)
    U = random_state((4, 5, 6), seed=4)
    W = NR.cons2prim(U, np.float64(G14))
    c = np.sqrt(G14 * W[4] / W[0])
    w = [1 / 6, 2 / 5, 0.5 / 4]
    invdt = max(((np.abs(W[d + 1]) + c) / w[d]).max() for d in range(3))
    assert O.compute_dt(g, U, 0.3) == 0.3 / invdt
    Uz = U.copy()
    Uz[1:4] = 0.0
    Uz[4] = 0.0
    Uz[4] += 1e-300 * 0
    assert O.compute_dt(g, np.zeros_like(U) + np.array([1, 0, 0, 0, 0.0])[:, None, None, None], 0.3) \
        == float("inf")
