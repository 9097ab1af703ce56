"""Writes the golden fixtures of this directory from the CPU oracle.

The reference holds no numeric fixtures for this path (its tests pin invariants
and thresholds only), so these vectors come from the oracle, which is pinned
against closed forms, the exact Riemann solver and the reference's thresholds
(tests/test_oracle.py). They freeze the oracle's outputs so that the GPU tests
can compare with committed data and any later change of the oracle shows up.

Run from the repository root:  python3 tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle import oracle as O  # noqa: E402
from helpers import prim2cons, random_state  # noqa: E402


def face_flux_kat():
    rng = np.random.default_rng(20240601)
    WL, WR, ax = [], [], []
    for axis in (1, 2, 3):
        for shift in (3.0, -3.0, 0.4, -0.4, 0.0):
            for _ in range(4):
                l = np.array([rng.uniform(0.3, 3), *rng.uniform(-0.6, 0.6, 3), rng.uniform(0.2, 3)])
                r = np.array([rng.uniform(0.3, 3), *rng.uniform(-0.6, 0.6, 3), rng.uniform(0.2, 3)])
                l[axis] += shift
                r[axis] += shift
                WL.append(l)
                WR.append(r)
                ax.append(axis)
    # stationary contact and identical supersonic states
    WL.append([1.0, 0, 0, 0, 1.0]); WR.append([0.125, 0, 0, 0, 1.0]); ax.append(1)
    WL.append([1.0, 5.0, 0, 0, 1.0]); WR.append([1.0, 5.0, 0, 0, 1.0]); ax.append(1)
    WL = np.ascontiguousarray(WL, dtype=np.float64)
    WR = np.ascontiguousarray(WR, dtype=np.float64)
    ax = np.ascontiguousarray(ax, dtype=np.int32)
    F64 = O.face_flux(WL, WR, ax, 1.4)
    F32 = O.face_flux(WL.astype(np.float32), WR.astype(np.float32), ax, 1.4)
    np.savez(os.path.join(HERE, "face_flux_kat.npz"), WL=WL, WR=WR, axis=ax, gamma=1.4, F64=F64, F32=F32)


def sod400():
    n = 400
    x = (np.arange(n) + 0.5) / n
    W = np.stack([np.where(x < 0.5, 1.0, 0.125), 0 * x, 0 * x, 0 * x, np.where(x < 0.5, 1.0, 0.1)])
    U0 = np.ascontiguousarray(prim2cons(W, 1.4))
    g = O.make_grid((n,), ((0.0, 1.0),), "outflow", 1.4)
    U = U0.copy()
    t, k = O.evolve(g, U, 0.2, 0.4)
    np.savez(os.path.join(HERE, "sod400_t0.2.npz"), U0=U0, U=U, t=t, steps=k)


def periodic8():
    U0 = random_state((8, 8, 8), seed=8, vmax=0.4)
    g = O.make_grid((8, 8, 8), [(0.0, 1.0)] * 3, "periodic", 1.4)
    U = np.ascontiguousarray(U0.copy())
    dts = []
    for _ in range(3):
        dt = O.compute_dt(g, U, 0.3)
        dts.append(dt)
        O.step(g, U, dt)
    U32 = np.ascontiguousarray(U0.astype(np.float32))
    for _ in range(3):
        O.step(g, U32, O.compute_dt(g, U32, 0.3, floor64=True), floor64=True)
    np.savez(os.path.join(HERE, "periodic8_3steps.npz"), U0=U0, U=U, dts=np.array(dts), U32=U32)


def step_vjp_2d():
    U = random_state((6, 8), seed=9, vmax=0.4)
    g = O.make_grid((8, 6), [(0.0, 1.0)] * 2, [("outflow", "reflecting"), ("periodic", "periodic")], 1.4)
    dt = 0.5 * O.compute_dt(g, U, 0.4)
    w = np.random.default_rng(10).normal(size=U.shape)
    np.savez(os.path.join(HERE, "step_vjp_8x6.npz"), U=U, w=w, dt=dt, G=O.step_vjp(g, U, dt, w))


if __name__ == "__main__":
    face_flux_kat()
    sod400()
    periodic8()
    step_vjp_2d()
    print("written:", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))
