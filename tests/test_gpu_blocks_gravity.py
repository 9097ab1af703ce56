"""Self-gravity on block meshes on the GPU against the oracle's restatement
(oracle/blocks_gravity.c, pinned in test_blocks_gravity_oracle.py):

* the composite operator A·x bit for bit (same CSR rows, Float64 row sums in
  the same order) on 2D and 3D two-level meshes, Float64 and Float32;
* g = −∇φ of a given φ bit for bit;
* one SSP-RK2 step with the gravity source of a given φ bit for bit (exact
  profile), Float64 and the FP32 form;
* the CG solve to its tolerance (dot products reduce in a different order),
  and a few single-rate steps of vh_blocks_run (a Poisson solve before each
  step) against the oracle's loop to that tolerance;
* a subcycled run (one solve per root step, g held over the substeps) against
  the oracle's evolve! with gravity to that tolerance.
"""
import sys

import numpy as np
import pytest

from conftest import load_vespa
from helpers import prim2cons

V = load_vespa()
from vespa_jl_amd import amr  # noqa: E402
from vespa_jl_amd import blockmesh as BM  # noqa: E402

pytestmark = pytest.mark.gpu
G14 = 1.4


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def mesh_2d():
    qb = amr.QuadBlocks(4, 8, 1, bc=BM.REF)
    qb.refine({(0, 1, 1), (0, 2, 1), (0, 1, 2)})
    return qb.mesh()


def mesh_3d():
    # a level-0 8³ block with a 2:1 pair of level-1 8×8×4... blocks against its hi x face
    h0 = (1.0 / 8,) * 3
    return BM.BlockMesh(3, 1, h0, G14, [0, 1, 1, 1, 1], [[8, 8, 8], [8, 8, 8], [8, 8, 8], [8, 8, 8], [8, 8, 8]],
                        [[0, 0, 0], [16, 0, 0], [16, 8, 0], [16, 0, 8], [16, 8, 8]], [[BM.OUT] * 6] * 5)


def state(m, dtype=np.float64, seed=0):
    cols = []
    for b in range(m.nblocks):
        X, Y, Z = m.cell_centers(b)
        X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
        r2 = (X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2 if m.rank == 3 else X * X + Y * Y
        rho = 1 + 1.5 * np.exp(-r2 / 0.03)
        cols.append(np.stack([rho, 0.1 * np.sin(2 * np.pi * Y), 0.1 * np.cos(2 * np.pi * X),
                              0.05 * np.ones_like(rho), 1 + 0.2 * np.sin(2 * np.pi * X)]))
    return prim2cons(np.concatenate(cols, axis=1), G14).astype(dtype)


MESHES = {"2d": mesh_2d, "3d": mesh_3d}


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("which", ["2d", "3d"])
def test_operator_g_and_sourced_step_bitwise(O, which, dtype):
    _log(f"block gravity {which} {np.dtype(dtype).name}")
    m = MESHES[which]()
    bo = O.BlockOracle(m)
    U = state(m, dtype)
    bs = BM.BlockSimulation(m, dtype, cfl=0.4)
    try:
        bs.set_mode("exact")
        bs.upload(U)
        bs.enable_gravity(G=1.0)
        x = np.random.default_rng(1).normal(size=m.ncell).astype(dtype)
        assert np.array_equal(bs.apply_laplacian(x), bo.poisson_apply(x))
        phi = np.zeros(m.ncell, dtype=dtype)
        bo.poisson_solve(np.ascontiguousarray(U), 1.0, True, 2000, 1e-10 if dtype == np.float64 else 1e-5, phi)
        bs.phi = phi
        assert np.array_equal(bs.gravity_accel(), bo.grav_accel(phi))
        dt = 0.5 * bo.compute_dt_grav(U, 0.4, 1.0, floor64=dtype == np.float32)
        assert dt == 0.5 * bs.compute_dt()
        Uo = np.ascontiguousarray(U.copy())
        bo.step_grav(Uo, dt, phi, floor64=dtype == np.float32)
        bs.step(dt)
        assert np.array_equal(bs.state(), Uo)
    finally:
        bs.close()


@pytest.mark.parametrize("which", ["2d", "3d"])
def test_cg_solve_and_run_within_tolerance(O, which):
    _log(f"block gravity CG and run {which}")
    m = MESHES[which]()
    bo = O.BlockOracle(m)
    U = state(m)
    bs = BM.BlockSimulation(m, np.float64, cfl=0.4)
    try:
        bs.upload(U)
        bs.set_mode("exact")
        bs.enable_gravity(G=1.0, maxiter=2000, tol=1e-10)
        it_g, rr_g = bs.solve_poisson()
        phi_o = np.zeros(m.ncell)
        it_o, rr_o = bo.poisson_solve(U, 1.0, True, 2000, 1e-10, phi_o)
        assert rr_g <= 1e-10 and abs(it_g - it_o) <= max(2, 0.05 * it_o)
        phi_g = bs.phi
        assert np.abs(phi_g - phi_o).max() <= 1e-8 * np.abs(phi_o).max()
        # three single-rate steps with a solve before each
        bs.upload(U)
        bs.phi = np.zeros(m.ncell)
        bs.evolve(max_steps=3)
        Uo = np.ascontiguousarray(U.copy())
        phi = np.zeros(m.ncell)
        for _ in range(3):
            bo.poisson_solve(Uo, 1.0, True, 2000, 1e-10, phi)
            dt = bo.compute_dt_grav(Uo, 0.4, 1.0)
            bo.step_grav(Uo, dt, phi)
        Ug = bs.state()
        assert np.abs(Ug - Uo).max(axis=1).max() <= 1e-9 * np.abs(Uo).max()
        # gravity off again: the plain step
        bs.disable_gravity()
        bs.upload(U)
        Up = np.ascontiguousarray(U.copy())
        dt = bo.compute_dt(Up, 0.4)
        bo.step(Up, dt)
        bs.step(dt)
        assert np.array_equal(bs.state(), Up)
    finally:
        bs.close()


def test_subcycled_run_with_gravity_within_tolerance(O):
    """evolve!(...; subcycle=true) with gravity (src/driver.jl:581-605): one
    Poisson solve per root step, g held over every substep, the level's CFL with
    the free-fall term; the GPU run against the oracle's, same root and substep
    counts, states within the CG tolerance's reach."""
    _log("block gravity subcycled run 2d")
    qb = amr.QuadBlocks(4, 8, 1, bc=BM.REF)
    qb.refine({(0, 1, 1), (0, 2, 1), (0, 1, 2), (0, 2, 2)})
    m = qb.mesh()
    x, y = qb.cell_centers()
    U = np.zeros((5, m.ncell))
    U[0] = 1 + 2.0 * np.exp(-(x * x + y * y) / 0.02)
    U[4] = 1.0 / (G14 - 1)
    bo = O.BlockOracle(m)
    Uo, phio = np.ascontiguousarray(U.copy()), np.zeros(m.ncell)
    to, ko, _, _ = bo.evolve_grav(Uo, 0.03, 0.4, 1.0, phio, subcycle=True)
    bs = BM.BlockSimulation(m, np.float64, cfl=0.4)
    try:
        bs.upload(U)
        bs.set_mode("exact")
        bs.enable_gravity(G=1.0, maxiter=2000, tol=1e-10)
        tg, kg = bs.evolve(tfinal=0.03, subcycle=True)
        Ug = bs.state()
    finally:
        bs.close()
    assert kg == ko and abs(tg - to) <= 1e-14
    assert np.abs(Ug - Uo).max() <= 1e-9 * np.abs(Uo).max()
