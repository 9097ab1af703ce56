"""Pins of the MUSCL oracle (CPU): PPMKernels' unsplit SSP-RK2 MUSCL step
(lib/PPMKernels/src/muscl.jl, lib/PPMKernels/src/muscl_grid.jl).

The reference certifies its flux line against the live Enzo solver (EnzoLib,
absent here). The restatement is pinned instead by:
  * known answers of the flux line: linear data are reconstructed exactly and
    the HLL flux of equal states is the Euler flux (closed form); a local
    extremum gets a zero slope; the θ-limited slope picks θ·(smaller jump);
    supersonic flow takes the upwind flux;
  * the reference's own driver checks (test/test_muscl_grid.jl): a 3D step on a
    state varying along one axis equals a 1D RK2 built from the flux line
    (x, y and z, exercising the cyclic momentum roles), conservation to
    round-off, and recorded fluxes that reproduce the update;
  * Sod against the exact Riemann solution, with first-order L1 convergence;
  * fill_periodic! against numpy's wrap padding.
"""
import numpy as np
import pytest

G14 = 1.4


def euler_flux(r, e, u, v, w, g=G14):
    """Closed-form Euler flux along the normal u (ρ, eint, u, v, w)."""
    p = (g - 1) * r * e
    E = r * (e + 0.5 * (u * u + v * v + w * w))
    return np.array([r * u, r * u * u + p, r * u * v, r * u * w, (E + p) * u])


def test_linear_data_reconstructed_exactly_and_hll_of_equal_states(O):
    """Linear profiles: θ ≥ 1 keeps the full slope, so L = R = the face value
    and HLL returns the Euler flux of that state."""
    n, ng = 16, 3
    x = np.arange(n, dtype=float)
    r, e, u, v, w = 1.0 + 0.01 * x, 2.0 + 0.03 * x, 0.1 + 0.002 * x, -0.05 + 0.001 * x, 0.02 - 0.003 * x
    F = np.array(O.muscl_flux_line(r, e, u, v, w, n, ng))
    xf = np.arange(ng, n - ng + 1) - 0.5                   # face positions
    lin = lambda a0, a1: a0 + a1 * xf
    ref = euler_flux(lin(1.0, 0.01), lin(2.0, 0.03), lin(0.1, 0.002), lin(-0.05, 0.001), lin(0.02, -0.003))
    np.testing.assert_allclose(F, ref, rtol=2e-15, atol=1e-17)


def _supersonic(rho, e=1.0, u=5.0):
    n = rho.size
    return rho, np.full(n, e), np.full(n, u), np.zeros(n), np.zeros(n)


def test_extremum_and_theta_limited_slope_upwind(O):
    """Supersonic to the right (am = 0): the flux is the Euler flux of the L state,
    the right-face value of the upwind cell. At a local maximum the slope is 0;
    with jumps 1 and 3 the slope is min(θ·1, θ·3, 2) = 1.5 (θ = 1.5), or 1 (θ = 1)."""
    n, ng = 12, 3
    rho = np.full(n, 10.0)
    rho[5:] = [11, 14, 20, 20, 20, 20, 20]                  # cell 5: jumps 1 (left) and 3 (right)
    rho[8] = 25.0                                           # cell 8: local maximum
    for th, s5 in ((1.5, 1.5), (1.0, 1.0)):
        fd = O.muscl_flux_line(*_supersonic(rho), n, ng, theta=th)[0]
        # face index f (0-based) sits between cells ng-1+f and ng+f; L = right face of ng-1+f
        np.testing.assert_allclose(fd[5 - ng + 1], (11 + 0.5 * s5) * 5.0, rtol=2e-16)
        np.testing.assert_allclose(fd[8 - ng + 1], 25.0 * 5.0, rtol=2e-16)
    # leftward supersonic: the R state (left face of the downwind cell) carries the flux
    r, e, u, v, w = _supersonic(rho, u=-5.0)
    fd = O.muscl_flux_line(r, e, u, v, w, n, ng)[0]
    np.testing.assert_allclose(fd[5 - ng], (11 - 0.5 * 1.5) * -5.0, rtol=2e-16)


def test_enzo_sign_of_zero_and_density_floor(O):
    """sign(0) = −1 makes the slope of an all-zero difference −0, so a −0 velocity
    keeps its sign through reconstruction and the transverse momentum flux is −0;
    reconstructed densities are floored at small_rho."""
    n, ng = 10, 3
    r, e, u, v, w = _supersonic(np.linspace(1.0, 2.0, n))
    w = np.full(n, -0.0)
    fs3 = O.muscl_flux_line(r, e, u, v, w, n, ng)[3]
    assert np.all(fs3 == 0) and np.all(np.signbit(fs3))
    fd = O.muscl_flux_line(*_supersonic(np.full(n, 1e-12)), n, ng, small_rho=1e-10)[0]
    np.testing.assert_allclose(fd, 1e-10 * 5.0, rtol=2e-16)


def _rk2_line_ref(O, D, Sn, St1, St2, Tau, na, ng, dt, dx, gamma, theta, grid_order=None):
    """1D SSP-RK2 from the flux line (test/test_muscl_grid.jl:20-44). grid_order:
    the grid axis of each velocity slot (normal, t1, t2), to sum |v|² in grid
    order as the 3D driver's cons2prim does (then the result is bit-exact)."""
    active = na - 2 * ng
    dtdx = dt / dx

    def Lop(D, Sn, St1, St2, Tau):
        vx, vy, vz = Sn / D, St1 / D, St2 / D
        if grid_order is None:
            eint = Tau / D - 0.5 * (vx ** 2 + vy ** 2 + vz ** 2)
        else:
            g = [None] * 3
            for q, a in enumerate(grid_order):
                g[a] = (vx, vy, vz)[q]
            eint = Tau / D - 0.5 * (g[0] * g[0] + g[1] * g[1] + g[2] * g[2])
        F = O.muscl_flux_line(D.copy(), eint, vx, vy, vz, na, ng, gamma=gamma, theta=theta)
        d = [np.zeros(na) for _ in range(5)]
        for m in range(active):
            c = ng + m
            for f in range(5):
                d[f][c] = -(F[f][m + 1] - F[f][m]) * dtdx
        return d

    U0 = [a.copy() for a in (D, Sn, St1, St2, Tau)]
    d1 = Lop(*U0)
    U1 = [U0[i] + d1[i] for i in range(5)]
    d2 = Lop(*U1)
    return [0.5 * (U0[i] + U1[i] + d2[i]) for i in range(5)]


@pytest.mark.parametrize("varying_transverse", [False, True])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_3d_step_reduces_to_1d_rk2(O, axis, varying_transverse):
    """test/test_muscl_grid.jl:46-115, same profiles, tolerance of its layer A;
    then a cold, fast variant whose transverse velocities vary along the line
    (still uniform across it): there eint is small beside ½|v|², so the order in
    which |v|² is summed shows in the result and pins the (normal, t1, t2) roles."""
    ng, gamma, theta, dt, dx = 3, 1.4, 1.5, 0.02, 1.0
    na = 22
    odims = [(na, 8, 7), (7, na, 8), (8, 7, na)][axis]
    m = np.arange(1, na + 1)
    t = np.tanh((m - 11.0) * 0.8)
    rho1, pr1, vn1 = 0.6 - 0.35 * t, 0.6 - 0.45 * t, 0.15 - 0.10 * t
    vt1, vt2 = np.full(na, 0.05), np.full(na, -0.03)
    if varying_transverse:
        pr1 = 1e-3 * pr1
        vn1 = 4.0 * vn1
        vt1 = 0.3 + 0.13 * np.sin(m)
        vt2 = -0.2 + 0.07 * np.cos(1.7 * m)
    eint1 = pr1 / ((gamma - 1) * rho1)
    etot1 = eint1 + 0.5 * (vn1 ** 2 + vt1 ** 2 + vt2 ** 2)
    line = [rho1, rho1 * vn1, rho1 * vt1, rho1 * vt2, rho1 * etot1]
    ref = _rk2_line_ref(O, *line, na=na, ng=ng, dt=dt, dx=dx, gamma=gamma, theta=theta)
    nx, ny, nz = odims
    shape = (nz, ny, nx)
    roles = [(1, 2, 3), (2, 3, 1), (3, 1, 2)][axis]        # grid slots of (normal, t1, t2)
    U = np.zeros((5, nx * ny * nz))
    bshape = [1, 1, 1]
    bshape[2 - axis] = na

    def bcast(a):
        return np.broadcast_to(a.reshape(bshape), shape).ravel()
    U[0] = bcast(line[0])
    U[4] = bcast(line[4])
    for q, slot in enumerate(roles):
        U[slot] = bcast(line[1 + q])
    O.muscl_step_3d(U, odims, ng, dt, gamma, theta, dx)
    got = U.reshape((5,) + shape)
    sel = [4, 4, 4]
    sel[2 - axis] = slice(None)
    exact = _rk2_line_ref(O, *line, na=na, ng=ng, dt=dt, dx=dx, gamma=gamma, theta=theta,
                          grid_order=[axis, (axis + 1) % 3, (axis + 2) % 3])
    for name, f in (("D", 0), ("Sn", roles[0]), ("St1", roles[1]), ("St2", roles[2]), ("Tau", 4)):
        q = {"D": 0, "Sn": 1, "St1": 2, "St2": 3, "Tau": 4}[name]
        g = got[f][tuple(sel)]
        np.testing.assert_allclose(g, ref[q], rtol=1e-12, atol=1e-14, err_msg=name)
        # with |v|² summed in grid order the transverse sweeps are exact no-ops:
        # bit-exact, which also fixes the cyclic (normal, t1, t2) roles
        assert np.array_equal(g, exact[q]), name


def _blob(m, ng_center, scale, amp_r=0.4, amp_p=0.3):
    c = (np.arange(1, m + 1) - ng_center) / scale
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    b = np.exp(-(X ** 2 + Y ** 2 + Z ** 2))
    rho = 1.0 + amp_r * b
    pr = 0.6 + amp_p * b
    vx, vy, vz = 0.10 * X * b, 0.08 * Y * b, -0.06 * Z * b
    etot = pr / ((G14 - 1) * rho) + 0.5 * (vx ** 2 + vy ** 2 + vz ** 2)
    return np.stack([rho, rho * vx, rho * vy, rho * vz, rho * etot]).reshape(5, -1).copy()


def _total(f, dims, ng, dx):
    nx, ny, nz = dims
    a = f.reshape(nz, ny, nx)[ng:nz - ng, ng:ny - ng, ng:nx - ng]
    return a.sum() * dx ** 3


def test_conservation(O):
    """test/test_muscl_grid.jl:117-147: three steps of a contained blob."""
    m, ng, dt, dx = 26, 3, 0.02, 1.0
    dims = (m, m, m)
    U = _blob(m, 13.5, 2.2)
    t0 = [_total(U[f], dims, ng, dx) for f in range(5)]
    for _ in range(3):
        O.muscl_step_3d(U, dims, ng, dt, G14, 1.5, dx)
    t1 = [_total(U[f], dims, ng, dx) for f in range(5)]
    assert abs(t1[0] - t0[0]) / abs(t0[0]) < 1e-12
    for f in (1, 2, 3):
        assert abs(t1[f] - t0[f]) < 1e-11
    assert abs(t1[4] - t0[4]) / abs(t0[4]) < 1e-12


def _rec_state(n=16, ng=3):
    nb = n + 2 * ng
    c = (np.arange(1, nb + 1) - ng - 0.5) / n
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    rho = 1.0 + 0.3 * np.sin(np.pi * 2 * X) * np.cos(np.pi * 2 * Y)
    pr = 0.6 + 0.2 * np.cos(np.pi * 2 * Z)
    vx, vy, vz = 0.2 * np.sin(np.pi * 2 * Y), 0.15 * np.cos(np.pi * 2 * Z), -0.1 * np.sin(np.pi * 2 * X)
    etot = pr / ((G14 - 1) * rho) + 0.5 * (vx ** 2 + vy ** 2 + vz ** 2)
    return np.stack([rho, rho * vx, rho * vy, rho * vz, rho * etot]).reshape(5, -1).copy(), (nb, nb, nb)


def test_flux_recording_reproduces_the_update(O):
    """test/test_muscl_grid.jl:663-696: ΔU = −Σ_a (frec[a][c+ê_a] − frec[a][c])·dt/dx
    to round-off on the active cells, and recording leaves the step bit-identical."""
    n, ng, dt = 16, 3, 0.008
    dx = 1.0 / n
    U, dims = _rec_state(n, ng)
    U0 = U.copy()
    frec = O.muscl_step_3d(U, dims, ng, dt, G14, 1.5, dx, record=True)
    nx, ny, nz = dims
    ea = (1, nx, nx * ny)
    idx = np.arange(U.shape[1]).reshape(nz, ny, nx)[ng:nz - ng, ng:ny - ng, ng:nx - ng].ravel()
    for f in range(5):
        div = np.zeros(idx.size)
        for a in range(3):
            div += (frec[a, f, idx + ea[a]] - frec[a, f, idx]) * (dt / dx)
        assert np.abs((U[f, idx] - U0[f, idx]) + div).max() < 1e-13
    U2 = U0.copy()
    O.muscl_step_3d(U2, dims, ng, dt, G14, 1.5, dx)
    assert np.array_equal(U2, U)


def _sod_l1(O, n):
    ng = 3
    dims = (n + 2 * ng, 2 * ng + 1, 2 * ng + 1)
    dx = 1.0 / n
    x = (np.arange(dims[0]) - ng + 0.5) * dx
    left = x < 0.5
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    e = p / ((G14 - 1) * rho)
    U = np.zeros((5,) + dims[::-1])
    U[0] = rho
    U[4] = rho * e
    U = U.reshape(5, -1).copy()
    t = 0.0
    while t < 0.2 - 1e-12:
        dt = min(0.4 * dx / O.muscl_max_wavespeed(U, G14), 0.2 - t)
        O.muscl_step_3d(U, dims, ng, dt, G14, 1.5, dx)
        t += dt
    r = U[0].reshape(dims[::-1])[ng, ng, ng:-ng]
    xs = x[ng:-ng]
    ex = np.array([O.exact_riemann_sample((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), G14, (xi - 0.5) / 0.2)[0]
                   for xi in xs])
    return np.abs(r - ex).mean()


def test_sod_against_exact_solution(O):
    e200, e400 = _sod_l1(O, 200), _sod_l1(O, 400)
    assert e200 < 1.2e-2
    assert e400 / e200 < 0.75


def test_internal_energy_floor(O):
    """eint = τ/D − ½|v|² is floored at small_rho before reconstruction: a state
    whose eint is negative everywhere still steps to finite values."""
    n, ng = 12, 3
    dims = (n, 2 * ng + 1, 2 * ng + 1)
    x = np.arange(n)
    rho = np.broadcast_to(1.0 + 0.5 * (x > 5), dims[::-1]).ravel()
    U = np.zeros((5, rho.size))
    U[0] = rho
    U[1] = rho * 2.0
    U[4] = rho * (0.5 * 4.0 - 1e-3)                     # eint = −1e-3
    O.muscl_step_3d(U, dims, ng, 1e-3, G14, 1.5, 1.0)
    assert np.isfinite(U).all()


def test_fill_periodic_and_max_wavespeed(O):
    rng = np.random.default_rng(0)
    dims, ng = (9, 8, 10), 2
    f = rng.normal(size=dims[::-1]).copy()
    inner = f[ng:-ng, ng:-ng, ng:-ng].copy()
    O.muscl_fill_periodic(f.reshape(-1), dims, ng)
    assert np.array_equal(f, np.pad(inner, ng, mode="wrap"))
    U = np.stack([rng.uniform(0.5, 2, 50), rng.normal(size=50), rng.normal(size=50), rng.normal(size=50),
                  rng.uniform(-0.5, 3, 50)])
    u, v, w = U[1] / U[0], U[2] / U[0], U[3] / U[0]
    eint = U[4] / U[0] - 0.5 * (u * u + v * v + w * w)
    ref = (np.sqrt(np.maximum(1.4 * 0.4 * eint, 0.0)) + np.maximum(np.abs(u), np.maximum(np.abs(v), np.abs(w)))).max()
    assert (eint < 0).any()
    assert O.muscl_max_wavespeed(U, 1.4) == pytest.approx(ref, rel=1e-15)


def register_ref(frec, dims, ng, dt, dx):
    """boundary_flux_register_ref (lib/PPMKernels/src/reflux_capture.jl:140-165):
    host loops over the full recorded fluxes, 1-based keys."""
    act = tuple(d - 2 * ng for d in dims)
    sc = float(dt) * float(dx) ** 2
    nx, ny, nz = dims
    fr = frec.reshape(3, 5, nz, ny, nx)

    def F(a, p, b1, b2, u1, u2):
        c = [0, 0, 0]
        c[a - 1], c[b1 - 1], c[b2 - 1] = p - 1, ng + u1 - 1, ng + u2 - 1
        return tuple(float(fr[a - 1, f, c[2], c[1], c[0]]) * sc for f in range(5))

    def key(a, k, b1, b2, u1, u2):
        return tuple(k if d == a else (u1 if d == b1 else u2) for d in (1, 2, 3))
    flux, interior = {}, {}
    for a in (1, 2, 3):
        b1, b2 = (2, 3) if a == 1 else ((1, 3) if a == 2 else (1, 2))
        for u2 in range(1, act[b2 - 1] + 1):
            for u1 in range(1, act[b1 - 1] + 1):
                for c in range(1, act[a - 1]):
                    interior[(a, key(a, c, b1, b2, u1, u2))] = F(a, ng + c + 1, b1, b2, u1, u2)
                flux[(a, "lo", key(a, 1, b1, b2, u1, u2))] = F(a, ng + 1, b1, b2, u1, u2)
                flux[(a, "hi", key(a, act[a - 1], b1, b2, u1, u2))] = F(a, ng + act[a - 1] + 1, b1, b2, u1, u2)
    return flux, interior, act


def test_boundary_flux_register_conservation_identity(O):
    """test/test_reflux_capture.jl:54-70 with the unsplit driver: the extensive change
    of each field over the active cells equals Σ lo-plane − Σ hi-plane fluxes; the
    interior register holds every active cell's hi face but the last."""
    n, ng, dt = 12, 3, 0.008
    dx = 1.0 / n
    U, dims = _rec_state(n, ng)
    U0 = U.copy()
    frec = O.muscl_step_3d(U, dims, ng, dt, G14, 1.5, dx, record=True)
    flux, interior, act = register_ref(frec, dims, ng, dt, dx)
    assert len(flux) == 3 * 2 * n * n and len(interior) == 3 * (n - 1) * n * n
    nx, ny, nz = dims
    sl = (slice(ng, nz - ng), slice(ng, ny - ng), slice(ng, nx - ng))
    for f in range(5):
        dQ = ((U[f] - U0[f]).reshape(nz, ny, nx)[sl]).sum() * dx ** 3
        slo = sum(v[f] for (a, side, _), v in flux.items() if side == "lo")
        shi = sum(v[f] for (a, side, _), v in flux.items() if side == "hi")
        assert dQ == pytest.approx(slo - shi, rel=1e-11, abs=1e-13)
    # an interior face is the lo face of the next cell: the x-face between active cells (1,1,1) and (2,1,1)
    i0 = ng + 1
    lin = (ng * ny + ng) * nx + i0
    assert interior[(1, (1, 1, 1))] == tuple(float(frec[0, f, lin]) * (dt * (dx * dx)) for f in range(5))


def _with_ge(U):
    rho = U[0]
    u, v, w = U[1] / rho, U[2] / rho, U[3] / rho
    eint = U[4] / rho - 0.5 * (u * u + v * v + w * w)
    return np.vstack([U, rho * eint]).copy()


def test_gas_energy_flux_upwinds(O):
    """The sixth flux is ρe·u of the upwind state: rightward supersonic takes the L
    state, leftward the R state (a velocity jump makes them differ)."""
    n, ng = 12, 3
    rho, e = np.full(n, 2.0), np.full(n, 0.5)
    z = np.zeros(n)
    for sgn in (1.0, -1.0):
        u = np.where(np.arange(n) < 6, 5.0, 6.0) * sgn     # jump at the face between cells 5 and 6
        F = O.muscl_flux_line(rho, e, u, z, z, n, ng, ge_flux=True)
        f = 6 - ng                                          # that face's index
        up = 5.0 if sgn > 0 else 6.0
        np.testing.assert_allclose(F[5][f], 2.0 * 0.5 * up * sgn, rtol=2e-16)
        np.testing.assert_allclose(F[0][f], 2.0 * up * sgn, rtol=2e-16)


def test_dual_energy_subsonic_is_inert(O):
    """test/test_muscl_grid.jl:550-593 (rk2 driver): with the gas energy on, a subsonic
    blob conserves mass, keeps ρe > 0, and its density is bit-identical to the run
    without it (the η₁ selection picks the total-energy eint)."""
    m, ng, dims = 26, 3, (26, 26, 26)
    A = _with_ge(_blob(m, 13.5, 2.2))
    B = _blob(m, 13.5, 2.2)
    m0 = _total(A[0], dims, ng, 1.0)
    for _ in range(3):
        O.muscl_step_3d(A, dims, ng, 0.02, G14, 1.5, 1.0)
        O.muscl_step_3d(B, dims, ng, 0.02, G14, 1.5, 1.0)
    assert abs(_total(A[0], dims, ng, 1.0) - m0) / m0 < 1e-11
    assert A[5].min() > 0
    assert np.array_equal(A[0], B[0])


def test_dual_energy_selection_and_sync(O):
    """A uniform state has no flux divergence, so one step only applies the end-of-step
    sync: a cold fast flow (γ(γ−1)·eint_tot ≤ η₁|v|²) takes eint from ρe, a warm slow
    one from τ; ρe and τ are then rewritten from the chosen eint."""
    ng, dims = 2, (5, 5, 5)
    n = 125

    def run(rho, vel, e_tot, e_dual):
        v2 = vel * vel
        U = np.stack([np.full(n, rho), np.full(n, rho * vel), np.zeros(n), np.zeros(n),
                      np.full(n, rho * (e_tot + 0.5 * v2)), np.full(n, rho * e_dual)])
        O.muscl_step_3d(U, dims, ng, 1e-3, G14, 1.5, 1.0)
        return U
    cold = run(2.0, 10.0, 1e-6, 5e-6)
    np.testing.assert_allclose(cold[5], 2.0 * 5e-6, rtol=1e-12)
    np.testing.assert_allclose(cold[4], 2.0 * (5e-6 + 50.0), rtol=1e-14)
    warm = run(2.0, 0.1, 2.0, 1.0)
    np.testing.assert_allclose(warm[5], 2.0 * 2.0, rtol=1e-12)
    np.testing.assert_allclose(warm[4], 2.0 * (2.0 + 0.005), rtol=1e-14)
    low = run(2.0, 0.1, 2.0, 5.0)                     # eint_tot ≤ ½·eint_dual: the advected one wins
    np.testing.assert_allclose(low[5], 2.0 * 5.0, rtol=1e-12)


def test_dual_energy_gas_energy_is_advected_conservatively(O):
    """A cold blob in a fast bulk flow takes eint from ρe everywhere (γ(γ−1)eint ≪
    η₁|v|²), so the sync keeps ρe and its change is the divergence of the recorded
    sixth flux ρe·u, cell by cell, to round-off."""
    m, ng, dims = 20, 3, (20, 20, 20)
    U = _blob(m, 10.5, 2.0, amp_r=0.5, amp_p=0.0)
    rho = U[0]
    vel = (1.0, 0.5, -0.25)
    eint = np.full(rho.size, 1e-5)
    for q in range(3):
        U[1 + q] = rho * vel[q]
    U[4] = rho * (eint + 0.5 * sum(v * v for v in vel))
    A = np.vstack([U, rho * eint]).copy()
    A0 = A.copy()
    dt, dx = 0.01, 1.0
    frec = O.muscl_step_3d(A, dims, ng, dt, G14, 1.5, dx, record=True)
    assert frec.shape == (3, 6, A.shape[1])
    nx, ny, nz = dims
    ea = (1, nx, nx * ny)
    idx = np.arange(A.shape[1]).reshape(nz, ny, nx)[ng:nz - ng, ng:ny - ng, ng:nx - ng].ravel()
    for f in (0, 5):
        div = sum((frec[a, f, idx + ea[a]] - frec[a, f, idx]) * (dt / dx) for a in range(3))
        assert np.abs((A[f, idx] - A0[f, idx]) + div).max() < 1e-13 * np.abs(A0[f]).max()
    assert A[5].min() > 0


# ---------------- bc! (ghost refill before each RK stage) ----------------

def _periodic_field(n, seed=5, nf=5):
    """A smooth periodic active state of n³ cells (5 or 6 fields)."""
    c = 2 * np.pi * (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    rho = 1.0 + 0.3 * np.sin(X + 0.2) * np.cos(Y) + 0.1 * np.sin(2 * Z + 0.5)
    vx, vy, vz = 0.3 * np.cos(Y + Z), 0.2 * np.sin(X - Z), 0.25 * np.cos(X + 0.3)
    pr = 0.7 + 0.2 * np.cos(X + Y + Z)
    eint = pr / ((G14 - 1) * rho)
    etot = eint + 0.5 * (vx ** 2 + vy ** 2 + vz ** 2)
    f = [rho, rho * vx, rho * vy, rho * vz, rho * etot] + ([rho * eint] if nf == 6 else [])
    return np.stack(f)


def _ghosted(act, ng):
    """(nf, n, n, n) active fields → (nf, N) with periodic ghost zones."""
    return np.stack([np.pad(a, ng, mode="wrap") for a in act]).reshape(act.shape[0], -1).copy()


def _periodic_bc(O, dims, ng):
    def bc(V, stage):
        for q in range(V.shape[0]):
            O.muscl_fill_periodic(V[q], dims, ng)
    return bc


@pytest.mark.parametrize("nf", [5, 6])
def test_bc_refills_before_each_stage_as_a_larger_periodic_domain(O, nf):
    """muscl_step_3d!(...; bc! = fill_periodic!) on an n³ periodic box equals, on
    the active cells, the step without bc! of a 3×3×3 tiling of the box with
    its own periodic ghosts (centre tile): a refill before each stage reproduces
    the periodic neighbourhood exactly; one before the first stage only would
    leave stale stage-1 ghosts, which the tiling does not have."""
    n, ng, dt, dx = 8, 3, 0.05, 0.5
    act = _periodic_field(n, nf=nf)
    U = _ghosted(act, ng)
    dims = (n + 2 * ng,) * 3
    O.muscl_step_3d_bc(U, dims, ng, dt, _periodic_bc(O, dims, ng), G14, 1.5, dx)
    big = np.stack([np.tile(a, (3, 3, 3)) for a in act])
    B = _ghosted(big, ng)
    bdims = (3 * n + 2 * ng,) * 3
    O.muscl_step_3d(B, bdims, ng, dt, G14, 1.5, dx)
    small = U.reshape((nf,) + dims[::-1])[:, ng:ng + n, ng:ng + n, ng:ng + n]
    centre = B.reshape((nf,) + bdims[::-1])[:, ng + n:ng + 2 * n, ng + n:ng + 2 * n, ng + n:ng + 2 * n]
    assert np.array_equal(small, centre)
    # a refill before the first stage only differs (the stage-1 ghosts are stale)
    U2 = _ghosted(act, ng)
    first_only = _periodic_bc(O, dims, ng)
    O.muscl_step_3d_bc(U2, dims, ng, dt, lambda V, s: first_only(V, s) if s == 0 else None, G14, 1.5, dx)
    assert not np.array_equal(U2.reshape((nf,) + dims[::-1])[:, ng:ng + n, ng:ng + n, ng:ng + n], centre)


def test_bc_noop_is_bc_nothing_and_ghosts_combine_the_unfilled_input(O):
    """A bc! that writes nothing reproduces bc! = nothing on every cell. With
    fill_periodic! as bc!, two inputs that differ only in their ghost zones give
    the same active cells (both are refilled before each stage) but different
    ghost cells: the step's combination uses the copy of the input taken before
    the first refill (muscl_grid.jl copies D0 before bcfill!)."""
    n, ng, dt, dx = 8, 3, 0.05, 0.5
    dims = (n + 2 * ng,) * 3
    U = _ghosted(_periodic_field(n), ng)
    ref = U.copy()
    O.muscl_step_3d(ref, dims, ng, dt, G14, 1.5, dx)
    a = U.copy()
    O.muscl_step_3d_bc(a, dims, ng, dt, lambda V, s: None, G14, 1.5, dx)
    assert np.array_equal(a, ref)
    ghost = np.ones(dims[::-1], dtype=bool)
    ghost[ng:-ng, ng:-ng, ng:-ng] = False
    ghost = ghost.reshape(-1)
    B = U.copy()
    B[:, ghost] *= np.random.default_rng(2).uniform(0.5, 1.5, size=(5, ghost.sum()))
    A = U.copy()
    for V in (A, B):
        O.muscl_step_3d_bc(V, dims, ng, dt, _periodic_bc(O, dims, ng), G14, 1.5, dx)
    assert np.array_equal(A[:, ~ghost], B[:, ~ghost])
    assert (A[:, ghost] != B[:, ghost]).mean() > 0.9


def test_bc_periodic_conserves_over_steps(O):
    n, ng, dt, dx = 12, 3, 0.05, 0.5
    dims = (n + 2 * ng,) * 3
    U = _ghosted(_periodic_field(n), ng)
    t0 = [_total(U[f], dims, ng, dx) for f in range(5)]
    for _ in range(3):
        O.muscl_step_3d_bc(U, dims, ng, dt, _periodic_bc(O, dims, ng), G14, 1.5, dx)
    t1 = [_total(U[f], dims, ng, dx) for f in range(5)]
    assert abs(t1[0] - t0[0]) <= 1e-13 * abs(t0[0])
    assert abs(t1[4] - t0[4]) <= 1e-13 * abs(t0[4])
    for f in (1, 2, 3):
        assert abs(t1[f] - t0[f]) <= 1e-12
