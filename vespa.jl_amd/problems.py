"""Problem specifications as code (the reference's problems/ directory and the
benchmark workloads of this path).

* ``sod_problem``         problems/sod_shock_tube.jl:223-234 (also in hydro.py)
* ``turbulence_state``    3D periodic smooth turbulence-like state, the
                          headline benchmark input (256³, γ = 5/3)
* ``kh_state``            2D periodic Kelvin–Helmholtz shear layer
* ``sedov_problem``       problems/sedov_blast.jl:258-279

The turbulence field is a random-phase sum over every integer wave vector with
1 ≤ |k|² ≤ kmax², one of each ±k pair (its first non-zero component positive),
cos(2π k·x + φ_k) with φ_k uniform in [0, 2π) from numpy's PCG64 seeded with
``seed`` (density) and ``seed + 1..3`` (velocity components). Because every mode
is a lattice mode, the sum is evaluated exactly as the real part of an inverse
FFT of a sparse spectrum: mode k at cell centre x_j = (j + ½)/n contributes
Re(e^{i(φ_k + π k·1/n)} e^{2πi k·j/n}).
"""
from __future__ import annotations

import numpy as np

from .hydro import Problem, prim2cons_array, sod_problem  # noqa: F401


def _lattice_modes(kmax: int, rank: int):
    r = np.arange(-kmax, kmax + 1)
    grids = np.meshgrid(*([r] * rank), indexing="ij")
    K = np.stack([g.ravel() for g in grids], axis=1)
    k2 = (K * K).sum(axis=1)
    K = K[(k2 >= 1) & (k2 <= kmax * kmax)]
    # keep one of each ±k pair: first non-zero component positive
    first = np.array([row[np.nonzero(row)[0][0]] for row in K])
    return K[first > 0]


def random_phase_field(n: int, rank: int, kmax: int, seed: int) -> np.ndarray:
    """Unit-RMS field Σ_k cos(2π k·x + φ_k) on an n^rank periodic grid, array
    axes ordered (z, y, x) so the last axis is x (the SoA column-major order)."""
    K = _lattice_modes(kmax, rank)
    rng = np.random.default_rng(seed)
    phi = rng.uniform(0.0, 2.0 * np.pi, len(K))
    A = np.zeros((n,) * rank, dtype=np.complex128)
    shift = np.pi * K.sum(axis=1) / n
    coef = np.exp(1j * (phi + shift))
    # axis order of A is reversed (z, y, x) ↔ k = (kx, ky, kz)
    idx = tuple((K[:, rank - 1 - a] % n) for a in range(rank))
    np.add.at(A, idx, coef)
    f = np.fft.ifftn(A).real * float(n) ** rank
    rms = np.sqrt(np.mean(f * f))
    return f / rms


def turbulence_state(n: int = 256, gamma: float = 5.0 / 3.0, seed: int = 20240601, kmax: int = 8,
                     dtype=np.float64):
    """Conserved state (5, n, n, n): ρ = 1 + 0.1·δ, each velocity component an
    independent unit-RMS field scaled to RMS 0.1·c_s with c_s = sqrt(γ·1/1), p = 1."""
    delta = random_phase_field(n, 3, kmax, seed)
    rho = 1.0 + 0.1 * delta
    cs = np.sqrt(gamma * 1.0 / 1.0)
    v = [0.1 * cs * random_phase_field(n, 3, kmax, seed + 1 + c) for c in range(3)]
    W = np.empty((5, n, n, n), dtype=np.float64)
    W[0] = rho
    W[1], W[2], W[3] = v
    W[4] = 1.0
    if not (W[0] > 0).all():
        raise ValueError("turbulence density went non-positive")
    return prim2cons_array(W.astype(dtype), gamma, dtype)


def kh_state(n: int = 512, gamma: float = 1.4, seed: int = 7, amp: float = 0.01, width: float = 0.025,
             kmax: int = 4, dtype=np.float64):
    """2D periodic Kelvin–Helmholtz layer on [0,1)²: vx = ±0.5 with tanh width
    `width` at y = 0.25 and 0.75, ρ = 1 inside the central band and 2 outside
    (a 1:2 density ratio), p = 2.5, plus seeded vy modes with |k| ≤ kmax of
    amplitude `amp`. Returns (5, n, n)."""
    y = (np.arange(n) + 0.5) / n
    prof = 0.5 * (np.tanh((y - 0.25) / width) - np.tanh((y - 0.75) / width))   # 1 inside band
    rho = (2.0 - prof)[:, None] * np.ones((1, n))
    vx = (0.5 * (2.0 * prof - 1.0))[:, None] * np.ones((1, n))
    vy = amp * random_phase_field(n, 2, kmax, seed)
    W = np.stack([rho, vx, vy, np.zeros((n, n)), np.full((n, n), 2.5)])
    return prim2cons_array(W.astype(dtype), gamma, dtype)


def sedov_problem(n: int = 64, E0: float = 1.0, rho0: float = 1.0, p0: float = 1e-5, gamma: float = 1.4,
                  r0=None, tfinal: float = 0.05) -> Problem:
    """2D Sedov blast on [0,1]², reflecting walls (problems/sedov_blast.jl:258-279)."""
    dx = 1.0 / n
    rhot = 2.0 * dx if r0 is None else float(r0)
    area = np.pi * rhot ** 2
    p_hot = (gamma - 1) * E0 / area

    def init(x, y, z):
        r = np.hypot(x - 0.5, y - 0.5)
        return (rho0, 0.0, 0.0, 0.0, np.where(r <= rhot, p_hot, float(p0)))

    return Problem("Sedov2D", (n, n), ((0.0, 1.0), (0.0, 1.0)), gamma, "reflecting", init, tfinal, 0.3)
