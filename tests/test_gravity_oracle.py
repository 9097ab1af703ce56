"""The gravity oracle (CPU) pinned by closed forms and by the reference's own
oracles (test/test_gravity_poisson.jl O1/O2/O2b, test/test_gravity_hydro.jl
S1/S2): Laplacian symmetry, null space and sign; the periodic Fourier-mode
solution at second order; g = −∇φ of an analytic φ; the source's sign (an
overdensity collapses) and the Jeans growth rate."""
import numpy as np
import pytest

from oracle import oracle as O

G53 = 5.0 / 3.0


def _grid(dims, bcs="periodic", L=1.0):
    return O.make_grid(dims, [(0.0, L)] * len(dims), bcs, G53)


def _centers(n, L=1.0):
    return (np.arange(n) + 0.5) * (L / n)


def test_laplacian_symmetric_nullspace_positive():
    """O1 (test/test_gravity_poisson.jl:20-50): ⟨x,Ay⟩ = ⟨Ax,y⟩, A·1 = 0, ⟨x,Ax⟩ > 0."""
    g = _grid((16,))
    x = _centers(16)
    a = np.sin(2 * np.pi * x) + 0.3 * np.cos(6 * np.pi * x)
    b = np.cos(2 * np.pi * x) - 0.7 * np.sin(4 * np.pi * x)
    Aa, Ab = O.poisson_apply(g, 1, a), O.poisson_apply(g, 1, b)
    assert abs(a @ Ab - b @ Aa) < 1e-12
    assert np.abs(O.poisson_apply(g, 1, np.ones(16))).max() < 1e-12
    s = np.sin(2 * np.pi * x)
    assert s @ O.poisson_apply(g, 1, s) > 0


def test_laplacian_stencil_closed_form():
    """(A x)_c = Σ_d (area_d / w_d)(2x_c − x_{c−e_d} − x_{c+e_d}) on a periodic 3D grid,
    and the Neumann rows drop the boundary faces."""
    dims = (6, 5, 4)
    g = O.make_grid(dims, [(0.0, 1.2), (0.0, 1.0), (0.0, 0.8)], "periodic", G53)
    rng = np.random.default_rng(3)
    x = rng.normal(size=dims[::-1])
    w = [1.2 / 6, 1.0 / 5, 0.8 / 4]
    area = [w[1] * w[2], w[0] * w[2], w[0] * w[1]]
    ref = np.zeros_like(x)
    for d, ax in enumerate((2, 1, 0)):
        ref += area[d] / w[d] * (2 * x - np.roll(x, 1, axis=ax) - np.roll(x, -1, axis=ax))
    got = O.poisson_apply(g, 7, x.ravel()).reshape(x.shape)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-13)
    # Neumann along x only: the two x boundary faces contribute nothing
    got_n = O.poisson_apply(g, 6, x.ravel()).reshape(x.shape)
    ref_n = ref.copy()
    c = area[0] / w[0]
    ref_n[:, :, 0] -= c * (x[:, :, 0] - x[:, :, -1])
    ref_n[:, :, -1] -= c * (x[:, :, -1] - x[:, :, 0])
    np.testing.assert_allclose(got_n, ref_n, rtol=1e-12, atol=1e-13)


def _fourier(n, m=1, A=0.1, G=1.0):
    g = _grid((n,))
    x = _centers(n)
    k = 2 * np.pi * m
    U = np.zeros((5, n))
    U[0] = 1.0 + A * np.cos(k * x)
    U[4] = 1.0 / (G53 - 1)
    phi = np.zeros(n)
    it, rr = O.poisson_solve(g, 1, U, G, True, 5000, 1e-12, phi)
    err = np.abs(phi - (-4 * np.pi * G * A / k ** 2) * np.cos(k * x)).mean()
    return err, it, rr


def test_fourier_mode_second_order():
    """O2 (test/test_gravity_poisson.jl:54-80): φ → −(4πGA/k²)cos(kx), ~4× drop per doubling."""
    e64, _, rr64 = _fourier(64)
    e128, _, _ = _fourier(128)
    assert rr64 < 1e-10
    assert e64 < 5e-3
    assert e64 / e128 > 3.5


def test_fourier_mode_2d():
    """O2b (test/test_gravity_poisson.jl:97-117): ρ = 1 + A cos kx cos ky."""
    n, A = 32, 0.1
    k = 2 * np.pi
    g = _grid((n, n))
    x = _centers(n)
    X, Y = np.meshgrid(x, x, indexing="xy")
    U = np.zeros((5, n, n))
    U[0] = 1 + A * np.cos(k * X) * np.cos(k * Y)
    U[4] = 1.0
    phi = np.zeros(n * n)
    it, rr = O.poisson_solve(g, 3, U, 1.0, True, 5000, 1e-11, phi)
    ex = -4 * np.pi * A / (2 * k ** 2) * np.cos(k * X) * np.cos(k * Y)
    assert rr < 1e-9
    assert np.abs(phi.reshape(n, n) - ex).mean() < 1e-2


def test_rhs_mean_projection():
    """b = −4πG(ρ − ρ̄)V sums to zero with the projection (the solvability condition)."""
    g = _grid((8, 6))
    rng = np.random.default_rng(1)
    U = np.zeros((5, 6, 8))
    U[0] = rng.uniform(0.5, 2.0, (6, 8))
    b = O.poisson_rhs(g, U, 2.0, True)
    assert abs(b.sum()) < 1e-12 * np.abs(b).sum()
    b0 = O.poisson_rhs(g, U, 2.0, False)
    V = (1 / 8) * (1 / 6)
    np.testing.assert_allclose(b0, -4 * np.pi * 2.0 * U[0].ravel() * V, rtol=1e-14)


def test_accel_closed_form():
    """g = −(φ(x+w) − φ(x−w))/(2w) for φ = cos kx: k·sin(kx)·sin(kw)/(kw); at a
    Neumann side the missing neighbour is φ_c at half the width."""
    n = 40
    g = _grid((n,))
    x = _centers(n)
    k, w = 2 * np.pi, 1.0 / n
    phi = np.cos(k * x)
    acc = O.grav_accel(g, 1, phi)
    np.testing.assert_allclose(acc[0], np.sin(k * x) * np.sin(k * w) / w, rtol=1e-12, atol=1e-13)
    assert np.all(acc[1:] == 0)
    gN = _grid((n,), "outflow")
    accN = O.grav_accel(gN, 0, phi)
    assert accN[0][0] == -(phi[1] - phi[0]) / (0.5 * w + w)
    assert accN[0][-1] == -(phi[-1] - phi[-2]) / (w + 0.5 * w)


def test_source_momentum_and_energy():
    """One step from rest in a fixed potential: the momentum change is the
    impulse −ρ∇φ·dt (to first order in dt) and with φ = 0 the step equals the plain step."""
    n = 64
    g = _grid((n,))
    x = _centers(n)
    U = np.zeros((5, n))
    U[0] = 1.0
    U[4] = 1.0 / (G53 - 1)
    phi = 0.01 * np.cos(2 * np.pi * x)
    U1 = U.copy()
    dt = 1e-3
    O.step_grav(g, U1, dt, 1, phi)
    gx = O.grav_accel(g, 1, phi)[0]
    np.testing.assert_allclose(U1[1], dt * U[0] * gx, rtol=1e-3, atol=1e-14)
    U2, U3 = U.copy(), U.copy()
    O.step_grav(g, U2, dt, 1, np.zeros(n))
    O.step(g, U3, dt)
    np.testing.assert_array_equal(U2, U3)


def test_compute_dt_free_fall():
    """Gas at rest with c → 0: dt = cfl·½/√(4πGρ_max)."""
    n = 16
    g = _grid((n,))
    U = np.zeros((5, n))
    U[0] = np.linspace(0.5, 2.0, n)
    U[4] = 1e-30
    dt = O.compute_dt_grav(g, U, 0.3, 1.0)
    assert dt == pytest.approx(0.3 * 0.5 / np.sqrt(4 * np.pi * 2.0), rel=1e-12)
    assert O.compute_dt_grav(g, U, 0.3, 0.0) == O.compute_dt(g, U, 0.3)


def _jeans_run(n, A, p0, k, tmax, every=1):
    g = _grid((n,))
    x = _centers(n)
    U = np.zeros((5, n))
    U[0] = 1.0 + A * np.cos(k * x)
    U[4] = p0 / (G53 - 1)
    phi = np.zeros(n)
    t, ts, amps = 0.0, [], []
    ck = np.cos(k * x)
    while t < tmax:
        O.poisson_solve(g, 1, U, 1.0, True, 500, 1e-8, phi)
        dt = min(O.compute_dt_grav(g, U, 0.3, 1.0), tmax - t)
        O.step_grav(g, U, dt, 1, phi)
        t += dt
        ts.append(t)
        amps.append(((U[0] - 1.0) * ck).sum() / (ck * ck).sum())
    return np.array(ts), np.array(amps)


def test_overdensity_collapses():
    """S1 (test/test_gravity_hydro.jl:33-52): the cos mode grows (attractive gravity)."""
    k = 2 * np.pi
    g = _grid((128,))
    x = _centers(128)
    U = np.zeros((5, 128))
    U[0] = 1.0 + 1e-3 * np.cos(k * x)
    U[4] = 0.06 / (G53 - 1)
    phi = np.zeros(128)
    a0 = ((U[0] - 1) * np.cos(k * x)).sum()
    for _ in range(40):
        O.poisson_solve(g, 1, U, 1.0, True, 500, 1e-8, phi)
        O.step_grav(g, U, O.compute_dt_grav(g, U, 0.3, 1.0), 1, phi)
    a1 = ((U[0] - 1) * np.cos(k * x)).sum()
    assert a1 / a0 > 1.05


def test_jeans_growth_rate():
    """S2 (test/test_gravity_hydro.jl:54-87): late-time growth rate within 8% of
    √(4πGρ₀ − c_s²k²)."""
    p0, k = 0.06, 2 * np.pi
    s_exact = np.sqrt(4 * np.pi - G53 * p0 * k ** 2)
    ts, amps = _jeans_run(256, 1e-4, p0, k, 2.0 / s_exact)
    h = len(ts) // 2
    s_meas = np.polyfit(ts[h:], np.log(np.abs(amps[h:]) / 1e-4), 1)[0]
    assert abs(s_meas - s_exact) / s_exact < 0.08
