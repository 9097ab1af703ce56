"""The coarse↔fine checks of test_blocks_mirror_accuracy.py on the GPU block path:

* Sod's problem turned by 45° with dynamic refinement (indicator every 4 steps,
  two levels) through the batched kernel and the device regrid, bit for bit the
  block oracle's AMR run at base 32²; at base 64² (finest 256²) the refined run
  beats the same-base uniform run in L1(ρ) against the exact solution and the
  uniform-fine run is within 1.5× of it (test/test_amr.jl:76-105 in 2D);
* one step on a two-level mesh and its mirror image: bit for bit the oracle's,
  hence the same local, bounded, first-order asymmetry.
"""
import sys

import numpy as np
import pytest

from conftest import load_vespa
from test_blocks_mirror_accuracy import (asymmetry_report, l1_rho, mirror_back, mirror_pair, oracle_amr)

V = load_vespa()
from vespa_jl_amd import amr  # noqa: E402
from vespa_jl_amd import blockmesh as BM  # noqa: E402

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_oblique_sod_amr_run_equals_oracle_bitwise(O):
    _log("oblique Sod AMR base 32^2, 2 levels, t=0.1: GPU vs oracle")
    qg = amr.QuadBlocks(4, 8, 2, bc=BM.OUT)
    r = amr.evolve_amr(qg, amr.oblique_sod_state(qg), 0.1, cfl=0.4, refine_above=0.05, every=4)
    qo = amr.QuadBlocks(4, 8, 2, bc=BM.OUT)
    Uo, t, steps = oracle_amr(O, qo, amr.oblique_sod_state(qo), 0.1)
    assert qg.blocks == qo.blocks
    assert r["steps"] == steps and r["t"] == t
    assert np.array_equal(r["U"], Uo)


def test_refined_oblique_sod_beats_base_and_nears_fine(O):
    base, tf = 8, 0.1
    l1 = {}
    for name, nroot, ml in (("base", base, 0), ("amr", base, 2), ("fine", base * 4, 0)):
        _log(f"oblique Sod {name}: {nroot}x{nroot} root blocks of 8x8, maxlevel {ml}")
        qb = amr.QuadBlocks(nroot, 8, ml, bc=BM.OUT)
        r = amr.evolve_amr(qb, amr.oblique_sod_state(qb), tf, cfl=0.4, refine_above=0.05, every=4)
        l1[name] = l1_rho(O, qb, r["U"][0], r["t"])
        if name == "amr":      # outflow boundaries: mass leaves the box, so no conservation check here
            assert max(b[0] for b in qb.blocks) == 2
    _log(f"  L1: {l1}")
    assert l1["amr"] < l1["base"]
    assert l1["fine"] <= 1.5 * l1["amr"]


@pytest.mark.parametrize("n", [8, 16])
def test_two_level_mirror_step_equals_oracle(O, n):
    M, U, Mm, Um, Ly = mirror_pair(n)
    bo, bom = O.BlockOracle(M), O.BlockOracle(Mm)
    dt = min(bo.compute_dt(U, 0.4), bom.compute_dt(Um, 0.4))
    out = []
    for m, u, b in ((M, U, bo), (Mm, Um, bom)):
        s = BM.BlockSimulation(m, np.float64, cfl=0.4)
        s.upload(u)
        s.step(dt)
        g = s.state()
        s.close()
        o = u.copy()
        b.step(o, dt)
        assert np.array_equal(g, o)
        out.append(g)
    r = asymmetry_report(M, U, out[0], mirror_back(M, Mm, out[1], Ly), dt)
    assert r["far_rel"] <= 1e-14 and 1e-6 < r["asym"] <= r["bound"]
