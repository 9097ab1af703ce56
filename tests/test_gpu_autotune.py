"""The launch autotuner on the GPU: it sweeps launch shapes that compute the
same values, so the state after tuning and stepping equals the untuned run bit
for bit; its cache is keyed by the workload and the device's compute-unit
count, and a second call applies the cached pick without sweeping."""
import json

import numpy as np
import pytest

from conftest import load_vespa
from helpers import smooth_state

V = load_vespa()
from vespa_jl_amd import autotune as AT  # noqa: E402
from vespa_jl_amd import blockmesh as BM  # noqa: E402

pytestmark = pytest.mark.gpu


def _grid_sim(U, dims):
    prob = V.Problem("tune", dims, ((0.0, 1.0),) * len(dims), 5.0 / 3.0, "periodic",
                     lambda x, y, z: (1.0, 0.0, 0.0, 0.0, 1.0), 1.0, 0.3)
    sim = V.Simulation(V.UniformMesh(dims, prob.domain), prob)
    sim.upload(U)
    return sim


def test_cu_count_is_positive():
    n = V.device_cu_count()
    assert n >= 1
    print(f"compute units: {n}", flush=True)


def test_tune_grid_keeps_values_and_caches(tmp_path, monkeypatch):
    monkeypatch.setenv("VESPAHIP_TUNE_CACHE", str(tmp_path / "tune.json"))
    dims = (48, 40, 36)
    U = smooth_state(dims, seed=3)
    ref = _grid_sim(U, dims)
    V.run_steps(ref, 3)
    Ur = ref.state()
    ref.close()
    sim = _grid_sim(U, dims)
    r1 = AT.tune_grid(sim, steps=2, zchunks=(8, 16))
    assert not r1["cached"] and len(r1["trials"]) == sim.tile_count() * 2
    V.run_steps(sim, 3)
    assert np.array_equal(sim.state(), Ur)
    r2 = AT.tune_grid(sim, steps=2)
    assert r2["cached"] and (r2["tile_variant"], r2["zchunk"]) == (r1["tile_variant"], r1["zchunk"])
    keys = list(json.load(open(tmp_path / "tune.json")))
    assert keys and all(k.endswith(f"cu{V.device_cu_count()}") for k in keys)
    sim.close()


def test_tune_blocks_keeps_values_and_sizes_by_cu(tmp_path, monkeypatch):
    monkeypatch.setenv("VESPAHIP_TUNE_CACHE", str(tmp_path / "tune.json"))
    m = BM.random_block_pairs(24, seed=9)
    U = BM.smooth_block_state(m, seed=4)
    ref = BM.BlockSimulation(m, np.float64, cfl=0.3)
    ref.upload(U)
    ref.evolve(max_steps=2)
    Ur = ref.state()
    ref.close()
    bs = BM.BlockSimulation(m, np.float64, cfl=0.3)
    bs.upload(U)
    r = AT.tune_blocks(bs, steps=2, max_tile_x=(16, 32), zchunks=(0, 8), wgs_per_cu=(0, 2))
    ncu = V.device_cu_count()
    assert {t["max_workgroups"] for t in r["trials"]} == {0, 2 * ncu}
    bs.upload(U)
    bs.evolve(max_steps=2)
    assert np.array_equal(bs.state(), Ur)
    bs.close()
