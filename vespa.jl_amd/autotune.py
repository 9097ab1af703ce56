"""Launch-shape autotuner for the stage kernels.

Sweeps the launch shapes the library exposes on the live workload, times a few
steps of each with HIP events (the same timed loop as the benchmark), keeps the
fastest, and caches the choice per workload signature in a JSON file
(``$VESPAHIP_TUNE_CACHE`` or ``~/.cache/vespa_jl_amd/autotune.json``) so later
runs apply it without sweeping.

* uniform grids: tile variant (``vh_grid_tile_count`` shapes) × z planes per
  workgroup;
* block meshes: widest tile along x × z planes per tile × persistent workgroups
  per compute unit.

Every candidate computes the same values (launch shape never changes the
arithmetic); the state advanced during the sweep is restored afterwards.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np


def cache_path() -> str:
    p = os.environ.get("VESPAHIP_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "vespa_jl_amd", "autotune.json")


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(d: dict) -> None:
    p = cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + ".tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass


def _device_sig() -> str:
    from ._lib import device_cu_count
    return f"cu{device_cu_count()}"


def _grid_key(sim) -> str:
    return (f"grid:{sim.mesh.rank}:{'x'.join(map(str, sim.mesh.dims))}:{sim.eltype.name}:{sim.mode}:"
            f"{_device_sig()}")


def tune_grid(sim, steps: int = 5, zchunks=(8, 16, 32), use_cache: bool = True) -> dict:
    """Pick the tile variant and z chunk of a uniform-grid Simulation."""
    key = _grid_key(sim)
    cache = _load() if use_cache else {}
    if key in cache:
        c = cache[key]
        sim.set_launch(c["tile_variant"], c["zchunk"])
        return dict(c, cached=True)
    U0, t0, n0 = sim.state(), sim.t, sim.step_count
    zs = zchunks if sim.mesh.rank == 3 else (0,)
    trials = []
    for v in range(sim.tile_count()):
        for z in zs:
            sim.set_launch(v, z)
            sim.upload(U0)
            r = sim.time_steps(steps)
            trials.append({"tile_variant": v, "zchunk": z, "step_ms": r["step_ms"], "tile": sim.launch()["tile"]})
    best = min(trials, key=lambda d: d["step_ms"])
    sim.set_launch(best["tile_variant"], best["zchunk"])
    sim.upload(U0)
    sim.t, sim.step_count = t0, n0
    out = {"tile_variant": best["tile_variant"], "zchunk": best["zchunk"], "step_ms": best["step_ms"],
           "tile": best["tile"], "trials": trials, "when": time.time()}
    cache[key] = out
    _store(cache)
    return dict(out, cached=False)


def _blocks_key(bsim) -> str:
    m = bsim.mesh
    sig = (int(m.nblocks), int(m.ncell), int(m.rank), int(np.bincount(m.dims[:, 0]).argmax()))
    return f"blocks:{':'.join(map(str, sig))}:{bsim.eltype.name}:{_device_sig()}"


def tune_blocks(bsim, steps: int = 3, max_tile_x=(16, 32), zchunks=(0, 8, 16), wgs_per_cu=(0, 4),
                use_cache: bool = True) -> dict:
    """Pick the tiling, z planes per tile and persistent workgroup count of a
    BlockSimulation (single-rate graph replays are timed)."""
    key = _blocks_key(bsim)
    cache = _load() if use_cache else {}
    if key in cache:
        c = cache[key]
        bsim.set_tiling(c["max_tile_x"])
        bsim.set_launch(zchunk=c["zchunk"], max_workgroups=c["max_workgroups"])
        return dict(c, cached=True)
    U0 = bsim.state()
    from ._lib import device_cu_count
    ncu = device_cu_count()
    trials = []
    for mt in max_tile_x:
        bsim.set_tiling(mt)
        for z in zchunks:
            for w in wgs_per_cu:
                mw = 0 if w == 0 else w * ncu
                bsim.set_launch(zchunk=z, max_workgroups=mw)
                bsim.upload(U0)
                r = bsim.time_steps(steps, method="graph", stage2=False)
                trials.append({"max_tile_x": mt, "zchunk": z, "max_workgroups": mw, "step_ms": r["step_ms"]})
    best = min(trials, key=lambda d: d["step_ms"])
    bsim.set_tiling(best["max_tile_x"])
    bsim.set_launch(zchunk=best["zchunk"], max_workgroups=best["max_workgroups"])
    bsim.upload(U0)
    out = dict(best, trials=trials, when=time.time())
    cache[key] = out
    _store(cache)
    return dict(out, cached=False)
