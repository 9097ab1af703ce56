"""Seeded input generators shared by the CPU and GPU tests."""
import numpy as np


def prim2cons(W, gamma, dtype=np.float64):
    """src/eos.jl:28-32 over a (5, ...) array (same expression order as the reference)."""
    W = np.asarray(W, dtype=dtype)
    T = np.dtype(dtype).type
    g = T(gamma)
    r, vx, vy, vz, p = W
    E = p / (g - T(1)) + (T(0.5) * r) * ((vx * vx + vy * vy) + vz * vz)
    return np.ascontiguousarray(np.stack([r, r * vx, r * vy, r * vz, E]))


def random_state(shape, seed, dtype=np.float64, gamma=1.4, vmax=0.5):
    """Random but physical primitive state → conserved array (5, *shape)."""
    rng = np.random.default_rng(seed)
    W = np.stack([rng.uniform(0.5, 2.0, shape), rng.uniform(-vmax, vmax, shape),
                  rng.uniform(-vmax, vmax, shape), rng.uniform(-vmax, vmax, shape),
                  rng.uniform(0.5, 2.0, shape)])
    return prim2cons(W.astype(dtype), gamma, dtype)


def smooth_state(dims, seed, gamma=5.0 / 3.0, kmax=3, dtype=np.float64):
    """Smooth periodic random-phase state on the unit box (small-size analogue of
    the 3D turbulence-like benchmark input)."""
    rng = np.random.default_rng(seed)
    rank = len(dims)
    axes = [(np.arange(n) + 0.5) / n for n in dims]
    grids = np.meshgrid(*reversed(axes), indexing="ij")
    X = list(reversed(grids))

    def field():
        f = np.zeros(tuple(reversed(dims)))
        for _ in range(6):
            k = rng.integers(-kmax, kmax + 1, rank)
            ph = rng.uniform(0, 2 * np.pi)
            f += np.cos(2 * np.pi * sum(k[d] * X[d] for d in range(rank)) + ph)
        return f / max(np.sqrt(np.mean(f * f)), 1e-12)

    rho = 1.0 + 0.1 * field()
    cs = np.sqrt(gamma)
    v = [0.1 * cs * field() for _ in range(3)]
    p = np.ones_like(rho)
    W = np.stack([rho, v[0], v[1], v[2], p]).astype(dtype)
    return prim2cons(W, gamma, dtype)


def random_faces(nf, seed, dtype=np.float64, kinds=("mixed",)):
    """Primitive face states covering subsonic, supersonic-left/right, strong
    jumps, stationary contacts and shear, plus axes 1..3."""
    rng = np.random.default_rng(seed)
    WL = np.stack([rng.uniform(0.1, 3, nf), rng.uniform(-3, 3, nf), rng.uniform(-3, 3, nf),
                   rng.uniform(-3, 3, nf), rng.uniform(0.01, 5, nf)], axis=1)
    WR = np.stack([rng.uniform(0.1, 3, nf), rng.uniform(-3, 3, nf), rng.uniform(-3, 3, nf),
                   rng.uniform(-3, 3, nf), rng.uniform(0.01, 5, nf)], axis=1)
    q = nf // 8
    # stationary contact: equal pressure, zero normal velocity
    WR[:q, 4] = WL[:q, 4]
    WL[:q, 1:4] = 0.0
    WR[:q, 1:4] = 0.0
    # identical states
    WR[q:2 * q] = WL[q:2 * q]
    # strongly supersonic to the right / left
    WL[2 * q:3 * q, 1:4] = 20.0
    WR[2 * q:3 * q, 1:4] = 20.0
    WL[3 * q:4 * q, 1:4] = -20.0
    WR[3 * q:4 * q, 1:4] = -20.0
    axis = rng.integers(1, 4, nf).astype(np.int32)
    return WL.astype(dtype), WR.astype(dtype), axis
