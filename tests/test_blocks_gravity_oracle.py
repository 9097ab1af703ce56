"""Pins of the oracle's self-gravity on block meshes (oracle/blocks_gravity.c):
the composite Poisson operator of src/gravity.jl:488-600 built from the block
face walk, CG, g = −∇φ and the gravity source of the stage.

* One block is the uniform restatement (gravity_impl.h, pinned in
  test_gravity_oracle.py by the reference's oracles): operator, CG iterates,
  g and a sourced step bit for bit, for periodic and Neumann φ boundaries.
* A mesh cut into same-level blocks is the uniform operator up to the order of
  each row's sum (round-off), with the same CG solution to its tolerance.
* Two-level meshes (2:1 faces): the operator is symmetric with zero row sums
  (the coarse↔fine coefficients are the fine area over ½(w_c + w_f), pushed to
  both rows), positive semi-definite, and its CSR rows list each coarse cell's
  2^(rank−1) fine neighbours; the solution for a smooth density mode is at least
  as accurate as the coarse uniform grid's and the uniform fine grid's is at
  least as accurate as it (the O2 mode of test/test_gravity_poisson.jl on a
  composite mesh); gravity pulls gas toward an overdensity (S1 of
  test/test_gravity_hydro.jl) with mass conserved to round-off.
"""
import numpy as np
import pytest

from conftest import load_vespa
from helpers import prim2cons

V = load_vespa()
from vespa_jl_amd import amr  # noqa: E402
from vespa_jl_amd import blockmesh as BM  # noqa: E402

G14 = 1.4


def one_block(n, rank, bc):
    dims = [n if d < rank else 1 for d in range(3)]
    h0 = tuple(1.0 / n for _ in range(rank))
    return BM.BlockMesh(rank, 0, h0, G14, [0], [dims], [[0, 0, 0]], [[bc] * 6])


def smooth_density_state(m, seed=0, amp=0.2):
    rng = np.random.default_rng(seed)
    cols = []
    for b in range(m.nblocks):
        X, Y, Z = m.cell_centers(b)
        X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
        rho = 1 + amp * np.cos(2 * np.pi * X + 0.3) * np.cos(2 * np.pi * Y + 0.1) * np.cos(2 * np.pi * Z)
        vx = 0.1 * np.sin(2 * np.pi * Y)
        vy = 0.1 * np.cos(2 * np.pi * X)
        cols.append(np.stack([rho, vx, vy, 0.05 * np.ones_like(rho), 1 + 0.1 * np.sin(2 * np.pi * X)]))
    return prim2cons(np.concatenate(cols, axis=1), G14)


@pytest.mark.parametrize("rank,n,bc", [(1, 24, BM.PER), (2, 12, BM.PER), (2, 10, BM.OUT), (3, 6, BM.PER),
                                       (3, 5, BM.REF)])
def test_one_block_is_the_uniform_restatement(O, rank, n, bc):
    m = one_block(n, rank, bc)
    bo = O.BlockOracle(m)
    names = {BM.PER: "periodic", BM.OUT: "outflow", BM.REF: "reflecting"}
    g = O.make_grid((n,) * rank, [(0.0, 1.0)] * rank, names[bc], G14)
    gper = (1 << rank) - 1 if bc == BM.PER else 0
    U = smooth_density_state(m)
    x = np.random.default_rng(1).normal(size=m.ncell)
    assert np.array_equal(bo.poisson_apply(x), O.poisson_apply(g, gper, x))
    pm = bc == BM.PER
    phi_b, phi_u = np.zeros(m.ncell), np.zeros(m.ncell)
    it_b, rr_b = bo.poisson_solve(U, 1.0, pm, 400, 1e-10, phi_b)
    it_u, rr_u = O.poisson_solve(g, gper, U.reshape((5,) + (n,) * rank), 1.0, pm, 400, 1e-10,
                                 phi_u.reshape((n,) * rank))
    assert (it_b, rr_b) == (it_u, rr_u) and np.array_equal(phi_b, phi_u)
    gb = bo.grav_accel(phi_b)
    gu = O.grav_accel(g, gper, phi_u.reshape((n,) * rank)).reshape(3, -1)
    assert np.array_equal(gb, gu)
    Ub, Uu = U.copy(), U.copy().reshape((5,) + (n,) * rank)
    dt = 0.5 * bo.compute_dt_grav(U, 0.4, 1.0)
    assert dt == 0.5 * O.compute_dt_grav(g, Uu, 0.4, 1.0)
    bo.step_grav(Ub, dt, phi_b)
    O.step_grav(g, Uu, dt, gper, phi_u.reshape((n,) * rank))
    assert np.array_equal(Ub, Uu.reshape(5, -1))


def test_same_level_split_is_the_uniform_operator(O):
    """A 2D 24×16 outflow grid cut into three same-level blocks: the same rows up
    to the order of each sum."""
    h = 1.0 / 24
    m = BM.BlockMesh(2, 0, (h, h), G14, [0, 0, 0], [[8, 16, 1], [10, 16, 1], [6, 16, 1]],
                     [[0, 0, 0], [8, 0, 0], [18, 0, 0]], [[BM.OUT] * 6] * 3)
    bo = O.BlockOracle(m)
    g = O.make_grid((24, 16), [(0.0, 1.0), (0.0, 16 * h)], "outflow", G14)
    # leaf order of the blocks vs the uniform column-major order
    idx = np.concatenate([(np.arange(16)[:, None] * 24 + (x0 + np.arange(nx))[None, :]).ravel()
                          for x0, nx in ((0, 8), (8, 10), (18, 6))])
    x = np.random.default_rng(3).normal(size=24 * 16)
    yb = bo.poisson_apply(x[idx])
    yu = O.poisson_apply(g, 0, x.reshape(16, 24)).ravel()[idx]
    assert np.abs(yb - yu).max() <= 1e-13 * np.abs(yu).max()
    rp, ci, cf, *_ = bo.poisson_csr()
    assert rp[-1] == 2 * (23 * 16 + 24 * 15)                    # every interior face, both rows


def two_level_quadtree(nroot=4, B=8, tags=((0, 1, 1), (0, 2, 1), (0, 1, 2)), bc=BM.OUT):
    qb = amr.QuadBlocks(nroot, B, 1, bc=bc)
    qb.refine(set(tags))
    return qb


def test_two_level_operator_structure(O):
    qb = two_level_quadtree()
    m = qb.mesh()
    bo = O.BlockOracle(m)
    rp, ci, cf, glo, ghi, gs = bo.poisson_csr()
    n = m.ncell
    x = np.random.default_rng(4).normal(size=n)
    y = np.random.default_rng(5).normal(size=n)
    Ax, Ay = bo.poisson_apply(x), bo.poisson_apply(y)
    assert abs(x @ Ay - y @ Ax) <= 1e-12 * abs(x @ Ay)          # symmetric
    assert np.array_equal(bo.poisson_apply(np.ones(n)), np.zeros(n))   # zero row sums (Neumann null space)
    assert x @ Ax > 0                                            # positive on non-constant vectors
    # the coarse↔fine coefficient: fine area over ½(w_c + w_f), both directions
    lv = np.repeat(m.level, m.counts)
    w0 = m.h0[0]
    for c in range(n):
        for k in range(rp[c], rp[c + 1]):
            j = ci[k]
            if lv[c] != lv[j]:
                assert cf[k] == (w0 / 2) / (0.5 * (w0 + w0 / 2))
    # a coarse cell next to a fine block lists both fine cells of each such face
    # (two faces at a corner of the refined region)
    nfine = {c: sum(lv[ci[k]] > lv[c] for k in range(rp[c], rp[c + 1])) for c in range(n) if lv[c] == 0}
    assert set(nfine.values()) == {0, 2, 4}


def _mode_error(O, qb, two_d, G=1.0, eps=0.1):
    """Max and volume-weighted L1 error of φ for ρ = 1 + ε·mode on [-0.5,0.5]² with
    Neumann walls, mode = cos(πx') (exact φ = −4πGε·mode/π²) or cos(πx')cos(πy')
    (exact −4πGε·mode/(2π²)), x' = x + ½, up to a constant."""
    m = qb.mesh()
    bo = O.BlockOracle(m)
    x, y = qb.cell_centers()
    mode = np.cos(np.pi * (x + 0.5)) * (np.cos(np.pi * (y + 0.5)) if two_d else 1.0)
    U = np.zeros((5, m.ncell))
    U[0] = 1 + eps * mode
    U[4] = 1.0
    phi = np.zeros(m.ncell)
    it, rr = bo.poisson_solve(U, G, True, 5000, 1e-12, phi)
    assert rr <= 1e-12
    exact = -4 * np.pi * G * eps * mode / ((2 if two_d else 1) * np.pi ** 2)
    vol = qb.volumes()
    exact -= (exact * vol).sum() / vol.sum()
    d = np.abs(phi - exact) / np.abs(exact).max()
    return float(d.max()), float((d * vol).sum() / vol.sum())


def test_composite_solution_accuracy(O):
    """Across a straight 2:1 interface normal to the mode's gradient (left half
    refined), the composite solution lies between the coarse and the fine
    uniform solutions in both norms. Across interfaces that the mode's gradient
    runs along, the two-point coarse↔fine flux of the reference's operator is
    only first-order consistent; the volume-weighted error still stays within a
    small multiple of the coarse grid's."""
    left = {(0, ix, iy) for ix in range(2) for iy in range(4)}
    for two_d in (False, True):
        ec = _mode_error(O, amr.QuadBlocks(4, 8, 0, bc=BM.OUT), two_d)
        ef = _mode_error(O, amr.QuadBlocks(8, 8, 0, bc=BM.OUT), two_d)
        assert ef[0] < 0.3 * ec[0] and ef[1] < 0.3 * ec[1]          # second order on uniform grids
        qb = amr.QuadBlocks(4, 8, 1, bc=BM.OUT)
        qb.refine(left)
        ea = _mode_error(O, qb, two_d)
        if not two_d:
            assert ef[0] < ea[0] < ec[0] and ef[1] < ea[1] < ec[1]
        for tags in (left, {(0, ix, iy) for ix in range(4) for iy in range(4) if (ix, iy) != (0, 0)},
                     {(0, 1, 1), (0, 2, 1), (0, 1, 2), (0, 2, 2)}):
            qb = amr.QuadBlocks(4, 8, 1, bc=BM.OUT)
            qb.refine(set(tags))
            assert _mode_error(O, qb, two_d)[1] < 5 * ec[1]


def test_gravity_pulls_toward_an_overdensity_on_two_levels(O):
    qb = two_level_quadtree(tags=((0, 1, 1), (0, 2, 1), (0, 1, 2), (0, 2, 2)), bc=BM.REF)   # closed box
    m = qb.mesh()
    bo = O.BlockOracle(m)
    x, y = qb.cell_centers()
    r2 = x * x + y * y
    U = np.zeros((5, m.ncell))
    U[0] = 1 + 2.0 * np.exp(-r2 / 0.02)
    U[4] = 1.0 / (G14 - 1)
    phi = np.zeros(m.ncell)
    bo.poisson_solve(U, 1.0, True, 5000, 1e-10, phi)
    g = bo.grav_accel(phi)
    rad = g[0] * x + g[1] * y
    assert (rad[r2 > 0.01] < 0).mean() > 0.99                  # g points inward
    U1 = np.ascontiguousarray(U.copy())
    dt = bo.compute_dt_grav(U, 0.4, 1.0)
    assert dt == bo.compute_dt(U, 0.4)                           # sound crossing binds at G = 1
    # at G = 100 the free-fall limiter √(4πGρ_max)/½ binds
    assert bo.compute_dt_grav(U, 0.4, 100.0) == 0.4 / (np.sqrt(4 * np.pi * 100.0 * U[0].max()) / 0.5)
    bo.step_grav(U1, dt, phi)
    mom_r = U1[1] * x + U1[2] * y
    assert mom_r[(r2 > 0.01) & (r2 < 0.09)].mean() < 0           # infall
    mass = lambda u: float((u[0] * qb.volumes()).sum())
    assert abs(mass(U1) - mass(U)) <= 1e-13 * mass(U)


def _blob_state(qb, amp=2.0):
    x, y = qb.cell_centers()
    U = np.zeros((5, qb.mesh().ncell))
    U[0] = 1 + amp * np.exp(-(x * x + y * y) / 0.02)
    U[4] = 1.0 / (G14 - 1)
    return U


def test_evolve_with_gravity_single_rate_and_subcycled(O):
    """evolve! with self-gravity (src/driver.jl:483-508, 581-605): a Poisson
    solve before each root step, the source in the stages of the stepped level,
    the free-fall limiter in each level's dt.

    * single rate: the loop solve → compute_dt → sourced step, bit for bit;
    * G = 0: the plain evolve!, single rate and subcycled, bit for bit (the hook
      adds exact zeros);
    * one level: subcycling takes one substep of the root dt, so it is the
      single-rate run bit for bit (the hook at level ≥ 0 is the hook at −1);
    * two levels in a closed box: mass conserved to round-off (the reflux
      keeps the coarse↔fine faces conservative), infall toward the blob, and
      within a few per cent of the single-rate run (g is held over the fine
      substeps: first order in the coupling)."""
    qb = two_level_quadtree(tags=((0, 1, 1), (0, 2, 1), (0, 1, 2), (0, 2, 2)), bc=BM.REF)
    m = qb.mesh()
    bo = O.BlockOracle(m)
    U0 = _blob_state(qb)
    # single rate against the explicit loop
    Ue, phie = np.ascontiguousarray(U0.copy()), np.zeros(m.ncell)
    t, k, it, rr = bo.evolve_grav(Ue, 1.0, 0.4, 1.0, phie, max_steps=3)
    assert k == 3 and it > 0 and rr <= 1e-10
    Ul, phil, tl = np.ascontiguousarray(U0.copy()), np.zeros(m.ncell), 0.0
    for _ in range(3):
        bo.poisson_solve(Ul, 1.0, True, 2000, 1e-10, phil)
        dt = bo.compute_dt_grav(Ul, 0.4, 1.0)
        bo.step_grav(Ul, dt, phil)
        tl += dt
    assert t == tl and np.array_equal(Ue, Ul) and np.array_equal(phie, phil)
    # G = 0 is the plain evolve!
    for sub in (False, True):
        Ug, phig = np.ascontiguousarray(U0.copy()), np.zeros(m.ncell)
        Up = np.ascontiguousarray(U0.copy())
        tg, kg, _, _ = bo.evolve_grav(Ug, 0.02, 0.4, 0.0, phig, subcycle=sub)
        tp, kp = bo.evolve(Up, 0.02, 0.4, subcycle=sub)
        assert (tg, kg) == (tp, kp) and np.array_equal(Ug, Up) and not phig.any()
    # one level: subcycled == single rate
    q1 = amr.QuadBlocks(4, 8, 1, bc=BM.REF)
    b1 = O.BlockOracle(q1.mesh())
    runs = []
    for sub in (False, True):
        U, phi = np.ascontiguousarray(_blob_state(q1)), np.zeros(q1.mesh().ncell)
        runs.append((b1.evolve_grav(U, 0.02, 0.4, 1.0, phi, subcycle=sub)[:2], U))
    assert runs[0][0] == runs[1][0] and np.array_equal(runs[0][1], runs[1][1])
    # two levels, subcycled
    tf = 0.05
    Us, phis = np.ascontiguousarray(U0.copy()), np.zeros(m.ncell)
    ts, ks, _, _ = bo.evolve_grav(Us, tf, 0.4, 1.0, phis, subcycle=True)
    Ur, phir = np.ascontiguousarray(U0.copy()), np.zeros(m.ncell)
    tr, kr, _, _ = bo.evolve_grav(Ur, tf, 0.4, 1.0, phir)
    assert abs(ts - tf) < 1e-12 and abs(tr - tf) < 1e-12 and ks < kr    # the root steps take the coarse dt
    vol = qb.volumes()
    mass = lambda u: float((u[0] * vol).sum())
    assert abs(mass(Us) - mass(U0)) <= 1e-13 * mass(U0)
    x, y = qb.cell_centers()
    r2 = x * x + y * y
    mom_r = Us[1] * x + Us[2] * y
    assert mom_r[(r2 > 0.01) & (r2 < 0.09)].mean() < 0
    d = float((np.abs(Us[0] - Ur[0]) * vol).sum() / (np.abs(Ur[0] - U0[0]) * vol).sum())
    assert d < 0.05, d
