#!/usr/bin/env python3
"""Benchmark of the ghost-free flux-divergence hot path on one MI355X.

Workload (default): 3D 256³ periodic smooth turbulence-like state, 5 conserved
fields SoA, γ = 5/3, PLM-minmod + HLLC + SSP-RK2 at CFL 0.3 with the CFL dt
reduced on the device every step, Float64. One "step" = one full SSP-RK2 step
over the whole grid (one zone-update per cell); inputs are resident in HBM
before the timed region starts.

Prints ONE JSON line on stdout (progress goes to stderr):
  value      = zone-updates/s (cells × steps / wall seconds of the timed region)
  roofline   = the fused stage-2 kernel: algorithmic bytes per launch (15 values
               per cell: read U¹ and Uⁿ, write Uⁿ⁺¹) ÷ its mean HIP-event duration
               inside the timed region, against 8 TB/s
  cpu_baseline = the oracle (C restatement of the RefMesh path, gather form,
               OpenMP on the host cores) on 2 steps of the same 256³ workload
Also reports the FP32 form (value, ratio to FP64, L1 error vs FP64).
"""
import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0


STAGE_SOURCES = ("common.h", "hydro.h", "stencil.h", "profiles.h", "stages.hip")


def stage_sources_hash() -> str:
    """sha256 over the sources the uniform stage kernels are compiled from."""
    import hashlib
    h = hashlib.sha256()
    for f in STAGE_SOURCES:
        with open(os.path.join(ROOT, "vespa.jl_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_package():
    if "vespa_jl_amd" in sys.modules:
        return sys.modules["vespa_jl_amd"]
    d = os.path.join(ROOT, "vespa.jl_amd")
    spec = importlib.util.spec_from_file_location(
        "vespa_jl_amd", os.path.join(d, "__init__.py"), submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["vespa_jl_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def make_sim(V, n, dtype, U, tile, zchunk, mode="fast"):
    prob = V.Problem("turbulence256", (n, n, n), ((0.0, 1.0),) * 3, 5.0 / 3.0, "periodic",
                     lambda x, y, z: (1.0, 0.0, 0.0, 0.0, 1.0), 1.0, 0.3)
    sim = V.Simulation(V.UniformMesh((n, n, n), prob.domain), prob, eltype=dtype)
    sim.set_mode(mode)
    if tile >= 0 or zchunk > 0:
        sim.set_launch(max(tile, 0), zchunk)
    sim.upload(U)
    return sim


def timed_run(V, sim, steps, warmup):
    """Warm up, then time exactly `steps` steps (wall clock, synchronised on both
    sides) with per-stage HIP events recorded on the library's stream."""
    if warmup > 0:
        V.run_steps(sim, warmup)
    sim.sync()
    t0 = time.perf_counter()
    ev = sim.time_steps(steps)          # synchronises the library's stream at the end
    sim.sync()
    wall = time.perf_counter() - t0
    return wall, ev


def host_cpu():
    """CPU model, logical CPUs of the host and those this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    return {"model": model, "logical_cpus": os.cpu_count(), "available_to_process": avail}


def baseline_threads():
    """Every host thread this process may use, capped by OMP_NUM_THREADS when the
    environment sets it (the GPU host allots 16 threads per GPU and exports
    OMP_NUM_THREADS=16)."""
    avail = host_cpu()["available_to_process"]
    env = os.environ.get("OMP_NUM_THREADS", "")
    return min(avail, int(env)) if env.isdigit() and int(env) > 0 else avail


def cpu_baseline(U, n, steps=2):
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    threads = baseline_threads()
    g = O.make_grid((n, n, n), [(0.0, 1.0)] * 3, "periodic", 5.0 / 3.0)
    Uc = np.ascontiguousarray(U.copy())
    work = np.empty(10 * n ** 3, dtype=Uc.dtype)
    t0 = time.perf_counter()
    for _ in range(steps):
        dt = O.compute_dt_omp(g, Uc, 0.3, threads=threads)
        O.step_omp(g, Uc, dt, threads=threads, work=work)
    el = time.perf_counter() - t0
    # the same oracle on one thread (RefMesh's serial loops), one step
    t0 = time.perf_counter()
    dt = O.compute_dt_omp(g, Uc, 0.3, threads=1)
    O.step_omp(g, Uc, dt, threads=1, work=work)
    el1 = time.perf_counter() - t0
    return {"value": n ** 3 * steps / el, "unit": "zone-updates/s", "cores": threads, "kind": "port",
            "cpu": host_cpu(),
            "sample": f"{steps} SSP-RK2 steps (CFL dt each) of the same {n}^3 FP64 workload, oracle in "
                      f"gather form with OpenMP on {threads} host threads, {el:.2f} s wall",
            "serial": {"value": n ** 3 / el1, "unit": "zone-updates/s", "cores": 1, "kind": "port",
                       "sample": f"1 further SSP-RK2 step of the same workload on one thread, {el1:.2f} s wall"}}


def blocks_leg(V, args):
    """Batched uneven AMR blocks: `args.blocks` blocks with edges drawn from
    {8, 16, 24, 32}, glued in random conforming and 2:1 pairs, smooth seeded
    state, single-rate SSP-RK2 with the device CFL dt. One persistent launch per
    RK stage replayed from a HIP graph (headline of the leg), the same launches
    enqueued eagerly, and one launch per block per stage (the comparison loop).
    Tile width, z planes per tile and the persistent workgroup count are picked
    by the autotuner first (timed graph replays of each candidate)."""
    import numpy as np
    from vespa_jl_amd import blockmesh as BM
    m = BM.random_block_pairs(args.blocks, seed=2024, rank=3, edges=(8, 16, 24, 32))
    log(f"[bench] blocks: {m.nblocks} blocks, {m.ncell} cells; building the state")
    U = BM.smooth_block_state(m, seed=1)
    res = {"workload": f"{m.nblocks} blocks (edges from {{8,16,24,32}}, seed 2024), {int((m.level == 1).sum())} "
                       f"level-1 blocks in 2:1 pairs, conforming pairs, smooth state rho,p in [0.5,2], "
                       f"|v|<=0.3, gamma=5/3, CFL 0.3 device dt, fast profile",
           "cells": m.ncell}
    U_end, dts = {}, []
    for dtype, key in ((np.float64, "f64"), (np.float32, "f32")):
        s = BM.BlockSimulation(m, dtype, cfl=0.3)
        s.set_mode(args.mode)
        s.upload(U.astype(dtype))
        tuned = None
        if not args.no_autotune:
            from vespa_jl_amd import autotune
            t0 = time.perf_counter()
            best = autotune.tune_blocks(s, use_cache=False)
            tuned = {"max_tile_x": best["max_tile_x"], "zchunk": best["zchunk"],
                     "max_workgroups": best["max_workgroups"], "candidates": len(best["trials"]),
                     "seconds": time.perf_counter() - t0}
            log(f"[bench] blocks {key} autotune: {tuned}")
        info = s.info()
        esz = np.dtype(dtype).itemsize
        r = {"tables": {k: info[k] for k in ("tiles", "link_faces", "register_cells")},
             "zchunk": info["zchunk"], "workgroups": info["workgroups"], "autotune": tuned}
        methods = args.block_methods.split(",")
        for method, nst in (("graph", args.block_steps), ("persistent", args.block_steps),
                            ("per_block", max(2, args.block_steps // 4))):
            if method not in methods or (key == "f32" and method != methods[0]):
                continue
            t = s.time_steps(nst, method=method, stage2=(method == "graph"))
            r[method] = {"ms_per_step": t["step_ms"], "zone_updates_per_s": m.ncell / (t["step_ms"] * 1e-3),
                         "steps": nst}
            if method == "graph":
                r["stage2_ms"] = t["stage2_ms"]
                r["stage2_GBps"] = 15 * esz * m.ncell / (t["stage2_ms"] * 1e-3) / 1e9
            log(f"[bench] blocks {key} {method}: {t['step_ms']:.3f} ms/step, "
                f"{m.ncell / (t['step_ms'] * 1e-3) / 1e9:.2f} G zone-updates/s")
        if key == "f64" and "graph" in r and "per_block" in r:
            r["speedup_graph_vs_per_block"] = r["per_block"]["ms_per_step"] / r["graph"]["ms_per_step"]
        # precision check: the same steps (FP64's CFL dt schedule) from the same state
        s.upload(U.astype(dtype))
        if key == "f64":
            dts = []
            for _ in range(10):
                dts.append(s.compute_dt())
                s.step(dts[-1])
        else:
            for dt in dts:
                s.step(dt)
        U_end[key] = s.state().astype(np.float64)
        s.close()
        res[key] = r
    den = np.abs(U_end["f64"]).mean(axis=1)
    res["f32"]["rel_L1_vs_fp64_per_field"] = dict(zip(
        ["density", "momentum_x", "momentum_y", "momentum_z", "energy"],
        (np.abs(U_end["f32"] - U_end["f64"]).mean(axis=1) / den).tolist()))
    res["f32"]["l1_after_steps"] = len(dts)
    log(f"[bench] blocks f32 rel L1 vs f64 after {len(dts)} steps: {res['f32']['rel_L1_vs_fp64_per_field']}")
    if not args.no_cpu_baseline:
        if ROOT not in sys.path:
            sys.path.insert(0, ROOT)
        from oracle import oracle as O
        comps = BM.components(m)
        sel = [b for g in comps[:300] for b in g]
        sub, cells = BM.submesh(m, sel)
        Us = np.ascontiguousarray(U[:, cells])
        bo = O.BlockOracle(sub)
        t0 = time.perf_counter()
        bo.step(Us, bo.compute_dt(Us, 0.3))
        el = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": sub.ncell / el, "unit": "zone-updates/s", "cores": 1, "kind": "port",
                               "sample": f"1 SSP-RK2 step (CFL dt) of the first 300 independent units "
                                         f"({len(sel)} blocks, {sub.ncell} cells) of the same block table, "
                                         f"serial block oracle, {el:.2f} s"}
        log(f"[bench] blocks CPU {sub.ncell / el / 1e6:.2f} M zone-updates/s")
    return res


def sod_leg(V, args):
    """Sod shock tube, 400 cells on [0,1], γ = 1.4, outflow, CFL 0.4, to t = 0.2
    (src/problem.jl:61-71): FP64 (exact profile) and the FP32 form on the GPU,
    L1(ρ) of each against the exact Riemann solution, the FP32 form's relative
    L1 per field against FP64, and the FP64 run against the oracle's."""
    import numpy as np
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from oracle import oracle as O
    prob = V.sod_problem(400)
    x = (np.arange(400) + 0.5) / 400
    exact = np.array([O.exact_riemann_sample((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4, (xi - 0.5) / 0.2)[0]
                      for xi in x])
    res = {"workload": "Sod 400 cells, gamma 1.4, outflow, CFL 0.4, t = 0.2"}
    ends = {}
    for dtype, key in ((np.float64, "f64"), (np.float32, "f32")):
        sim = V.Simulation(V.UniformMesh(prob.dims, prob.domain), prob, eltype=dtype)
        sim.set_mode("exact")
        U0 = sim.state().copy()
        t0 = time.perf_counter()
        V.evolve(sim)
        el = time.perf_counter() - t0
        U = sim.state().astype(np.float64)
        ends[key] = U
        res[key] = {"steps": sim.step_count, "t": sim.t, "wall_s": el,
                    "l1_rho_vs_exact": float(np.abs(U[0] - exact).mean())}
        if key == "f64":
            g = O.make_grid((400,), ((0.0, 1.0),), "outflow", 1.4)
            Uo = np.ascontiguousarray(U0.copy())
            to, ko = O.evolve(g, Uo, 0.2, 0.4)
            res[key]["oracle_bitwise"] = bool(np.array_equal(Uo, sim.state()) and ko == sim.step_count and
                                              to == sim.t)
        sim.close()
    den = np.abs(ends["f64"]).mean(axis=1)
    res["f32"]["rel_L1_vs_fp64_per_field"] = dict(zip(
        ["density", "momentum_x", "momentum_y", "momentum_z", "energy"],
        (np.abs(ends["f32"] - ends["f64"]).mean(axis=1) / np.maximum(den, 1e-300)).tolist()))
    log(f"[bench] Sod-400: L1 f64 {res['f64']['l1_rho_vs_exact']:.4e}, f32 {res['f32']['l1_rho_vs_exact']:.4e}, "
        f"oracle bitwise {res['f64']['oracle_bitwise']}")
    return res


def adjoint_leg(V, args):
    """Kelvin-Helmholtz 512^2 periodic FP64: 200 steps (device CFL dt), loss
    L = sum over cells of (rho_final - rho_target)^2 with the target the final
    density of the same run from a second seeded state; gradient w.r.t. the
    initial conserved fields by the hand-written adjoint with checkpoints every
    20 steps. Reports the backward/forward time ratio, device memory, central
    differences on random directions (eps = 1e-6) and the GPU gradient against
    the oracle's chained VJPs on a 128^2 version of the same rollout."""
    import numpy as np
    from vespa_jl_amd import adjoint as AD
    from vespa_jl_amd import problems as P
    n, nst = args.adjoint_n, args.adjoint_steps

    def sim_for(U):
        prob = V.Problem("kh", (n, n), ((0.0, 1.0),) * 2, 1.4, "periodic",
                         lambda x, y, z: (1.0, 0.0, 0.0, 0.0, 1.0), 1.0, 0.4)
        s = V.Simulation(V.UniformMesh((n, n), prob.domain), prob)
        s.upload(U)
        return s

    U0 = np.ascontiguousarray(P.kh_state(n, seed=7))
    st = sim_for(np.ascontiguousarray(P.kh_state(n, seed=8)))
    V.run_steps(st, nst)
    target = st.state()[0]
    st.close()
    sim = sim_for(U0)
    # plain forward timing (same steps, device dt, no checkpoints)
    sim.sync()
    t0 = time.perf_counter()
    V.run_steps(sim, nst)
    sim.sync()
    plain_ms = 1e3 * (time.perf_counter() - t0)
    sim.upload(U0)
    sim.t = 0.0
    ro = AD.Rollout(sim, nst, every=20)
    ro.forward()                      # warm
    sim.upload(U0)
    UF = ro.forward()
    w = np.zeros_like(UF)
    w[0] = 2.0 * (UF[0] - target)
    ro.backward(w)                    # warm
    G = ro.backward(w)
    info = ro.info()
    dts = info["dts"]
    loss0 = float(np.sum((UF[0] - target) ** 2))

    def loss_of(U):
        sim.upload(U)
        for dt in dts:
            V.step(sim, float(dt))
        return float(np.sum((sim.state()[0] - target) ** 2))

    rng = np.random.default_rng(2024)
    dirs = [rng.normal(size=U0.shape) for _ in range(max(args.fd_directions, 0))]
    dirs = [v / np.linalg.norm(v) for v in dirs]
    fd = [{"adjoint": 0.0, "fd": 0.0, "rel": 0.0}] if not dirs else []
    sweep = {}
    for eps in (1e-4, 1e-5, 1e-6, 1e-7):
        rels = []
        for v in dirs:
            ad = float(np.sum(G * v))
            f = (loss_of(U0 + eps * v) - loss_of(U0 - eps * v)) / (2 * eps)
            rels.append(abs(ad - f) / max(abs(f), 1e-300))
            if eps == 1e-6:
                fd.append({"adjoint": ad, "fd": f, "rel": rels[-1]})
        if rels:
            sweep[f"{eps:g}"] = {"max_rel": max(rels), "median_rel": float(np.median(rels))}
    ro.close()
    sim.close()
    res = {"workload": f"KH {n}^2 periodic, gamma=1.4, {nst} steps CFL 0.4, checkpoints every 20, FP64",
           "forward_ms": info["forward_ms"], "forward_plain_ms": plain_ms, "backward_ms": info["backward_ms"],
           "ratio_backward_over_forward": info["backward_ms"] / info["forward_ms"],
           "device_bytes_rollout": info["device_bytes"], "loss": loss0,
           "fd_eps": 1e-6, "fd_max_rel": max(d["rel"] for d in fd), "fd_median_rel": float(np.median([d["rel"] for d in fd])),
           "fd": fd, "fd_eps_sweep": sweep}
    log(f"[bench] adjoint: fwd {info['forward_ms']:.2f} ms, bwd {info['backward_ms']:.2f} ms, "
        f"ratio {res['ratio_backward_over_forward']:.2f}, FD max rel {res['fd_max_rel']:.2e}")
    # GPU gradient vs the oracle's chained VJPs (same schedule) at 128^2
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from oracle import oracle as O
    m = 128
    Us = np.ascontiguousarray(P.kh_state(m, seed=7))
    prob = V.Problem("kh", (m, m), ((0.0, 1.0),) * 2, 1.4, "periodic",
                     lambda x, y, z: (1.0, 0.0, 0.0, 0.0, 1.0), 1.0, 0.4)
    s2 = V.Simulation(V.UniformMesh((m, m), prob.domain), prob)
    s2.upload(Us)
    r2 = AD.Rollout(s2, nst, every=20)
    UF2 = r2.forward()
    w2 = np.zeros_like(UF2)
    w2[0] = 2.0 * (UF2[0] - np.ascontiguousarray(P.kh_state(m, seed=8))[0])
    G2 = r2.backward(w2)
    d2 = r2.info()["dts"]
    r2.close()
    s2.close()
    g = O.make_grid((m, m), [(0.0, 1.0)] * 2, "periodic", 1.4)
    states, U = [], Us.copy()
    t0 = time.perf_counter()
    for dt in d2:
        states.append(U.copy())
        O.step(g, U, float(dt))
    wr = w2.copy()
    for k in reversed(range(nst)):
        wr = O.step_vjp(g, states[k], float(d2[k]), wr)
    el = time.perf_counter() - t0
    res["vs_oracle_gradient"] = {"grid": [m, m], "steps": nst,
                                 "forward_bitwise": bool(np.array_equal(U, UF2)),
                                 "max_rel_diff": float(np.abs(G2 - wr).max() / np.abs(wr).max()),
                                 "oracle_seconds": el}
    log(f"[bench] adjoint vs oracle at {m}^2: {res['vs_oracle_gradient']['max_rel_diff']:.2e}")
    # the transpose identity <G, v> = <w, J v> with J v from the oracle's forward
    # mode (dual numbers, the reverse mode's branch rules) on the same schedule
    nd = O.jvp_max_directions()
    Vd = np.random.default_rng(11).normal(size=(nd,) + Us.shape)
    t0 = time.perf_counter()
    Un, JV = O.rollout_jvp(g, Us, d2, Vd, threads=baseline_threads())
    el = time.perf_counter() - t0
    lhs = Vd.reshape(nd, -1) @ G2.ravel()
    rhs = JV.reshape(nd, -1) @ w2.ravel()
    res["transpose_identity"] = {"grid": [m, m], "steps": nst, "directions": nd,
                                 "max_rel_gap": float(np.abs(lhs - rhs).max() / np.abs(rhs).max()),
                                 "forward_mode_states_bitwise": bool(np.array_equal(Un, UF2)),
                                 "oracle_seconds": el}
    log(f"[bench] adjoint transpose identity at {m}^2: {res['transpose_identity']['max_rel_gap']:.2e}")
    return res


def sedov_leg(V, args):
    """2D Sedov blast with dynamic block AMR (base 48^2 on [-0.5,0.5]^2, 8x8-cell
    blocks, up to 3 levels, p_amb 1e-5, E=1 in r<0.02), to t=0.05, FP64 and FP32.
    The regrids run on the device (indicator and remap kernels); their wall time
    per phase is reported beside the whole run's."""
    import numpy as np
    from vespa_jl_amd import amr
    out = {}
    # Untimed warm-up: whole runs of the same problem for at least half a second.
    # Measured right after the adjoint leg, the first ~0.1 s or more of host work
    # went 5x slower (table builds and the pure-Python tree edit alike): a
    # quarter-length warm-up did not absorb it, and once a whole run took 0.045 s
    # one run did not either (tables 10.3 ms against 4.9 ms).
    tw = time.perf_counter()
    while True:
        amr.run_sedov(tfinal=0.05, dtype=np.float64, mode=args.mode)
        if time.perf_counter() - tw >= 0.5:
            break
    for dtype, key in ((np.float64, "f64"), (np.float32, "f32")):
        t0 = time.perf_counter()
        r = amr.run_sedov(tfinal=0.05, dtype=dtype, snapshots=(0.0125, 0.025, 0.05), mode=args.mode)
        el = time.perf_counter() - t0
        rho = r["snapshots"][0.05][1]
        sc = rho.max()
        radii = {str(t): amr.shock_radius(v[1]) for t, v in r["snapshots"].items()}
        ts = np.array(sorted(r["snapshots"]))
        Rs = np.array([amr.shock_radius(r["snapshots"][t][1]) for t in ts])
        expo = float(np.polyfit(np.log(ts), np.log(Rs), 1)[0])
        out[key] = {"steps": r["steps"], "wall_s": el, "final_leaves": r["leaves"], "max_leaves": r["max_leaves"],
                    "mass_rel_change": float(abs(r["totals"][0] / r["totals0"][0] - 1)),
                    "energy_rel_change": float(abs(r["totals"][4] / r["totals0"][4] - 1)),
                    "symmetry_max_rel": float(max(np.abs(rho - rho[::-1, :]).max(), np.abs(rho - rho[:, ::-1]).max(),
                                                  np.abs(rho - rho.T).max()) / sc),
                    "shock_radius": radii, "radius_exponent": expo,
                    "indicator_calls": r["indicator_calls"], "regrids": r["regrids"],
                    "regrid_ms": {k: round(v, 3) for k, v in r["regrid_ms"].items()}}
        log(f"[bench] sedov {key}: {r['steps']} steps in {el:.3f} s, leaves {r['leaves']}, R ~ t^{expo:.3f}, "
            f"mass {out[key]['mass_rel_change']:.1e}, symmetry {out[key]['symmetry_max_rel']:.1e}, "
            f"{r['regrids']} regrids, regrid ms {out[key]['regrid_ms']}")
    return out


def muscl_turbulence_state(n, ng=4, nmodes=24, seed=1234, mach=0.3, gamma=1.4):
    """PPMKernels' turbulence benchmark state (lib/PPMKernels/bench/bench_turbulence.jl:31-57)
    on an (n + 2·ng)³ grid: ρ = 1, P = 1/γ (c_s = 1), a solenoidal sum of random
    Fourier modes |k_i| ≤ 3 with amplitude 1/|k| scaled to RMS Mach `mach`; our own
    seeded generator (numpy), the same construction."""
    import numpy as np
    rng = np.random.default_rng(seed)
    N = n + 2 * ng
    i = np.arange(N, dtype=np.float64)
    v = [np.zeros((N, N, N)) for _ in range(3)]
    for _ in range(nmodes):
        k = rng.integers(-3, 4, 3)
        if not k.any():
            continue
        k2 = float(k @ k)
        a = rng.standard_normal(3)
        a = a - (a @ k) / k2 * k
        ph = 2 * np.pi * rng.random()
        c = np.cos(2 * np.pi * (k[0] * i[None, None, :] + k[1] * i[None, :, None] + k[2] * i[:, None, None]) / N + ph)
        c *= 1.0 / np.sqrt(k2)
        for q in range(3):
            v[q] += a[q] * c
    vrms = np.sqrt(sum((x * x).mean() for x in v))
    v = [x * (mach / vrms) for x in v]
    rho = np.ones((N, N, N))
    ge = (1.0 / gamma) / (gamma - 1.0)
    etot = ge + 0.5 * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2)
    U = np.stack([rho, rho * v[0], rho * v[1], rho * v[2], rho * etot]).reshape(5, -1)
    return U, (N, N, N)


def muscl_leg(V, args):
    """PPMKernels unsplit RK2 MUSCL (PLM minmod-θ + HLL) on the reference's turbulence
    benchmark: 256³ active cells + 4 ghost zones, γ = 1.4, θ = 1.5, dx = 1/256,
    dt = 0.2·dx, Float32 and Float64, fused stage kernels in the exact (bit-identical)
    and the fast arithmetic profile.
    Mcell/s counts active cells per second, as the reference's table does; the
    CPU baseline is the oracle (one thread) on a 32³ + ghost sample."""
    import numpy as np
    from vespa_jl_amd import muscl as M
    n, ng = args.muscl_n, 4
    U64, dims = muscl_turbulence_state(n, ng)
    dx = 1.0 / n
    dt = 0.2 * dx
    ncell = int(np.prod(dims))
    out = {"workload": f"{n}^3 active + {ng} ghost zones per side, decaying-turbulence state, Mach 0.3",
           "published_reference": {"value_Mcell_s": 96.51, "hardware": "Apple Silicon, Metal, Float32",
                                   "source": "lib/PPMKernels/bench/RESULTS_turbulence.md:24"}}
    for dtype, key in ((np.float32, "f32"), (np.float64, "f64")):
        for mode in ("exact", "fast"):
            g = M.MusclGrid(dims, ng, 1.4, 1.5, dx, eltype=dtype)
            g.set_mode(mode)
            g.upload(U64.astype(dtype))
            r = g.time_steps(dt, args.muscl_steps)
            esz = np.dtype(dtype).itemsize
            k = key if mode == "exact" else key + "_fast"
            out[k] = {"Mcell_s": n ** 3 / (r["step_ms"] * 1e-3) / 1e6, "ms_per_step": r["step_ms"],
                      "stage2_ms": r["stage2_ms"],
                      "stage2_GBps": 15 * esz * ncell / (r["stage2_ms"] * 1e-3) / 1e9,
                      "speedup_vs_published": n ** 3 / (r["step_ms"] * 1e-3) / 1e6 / 96.51,
                      "finite": bool(np.isfinite(g.state()).all())}
            g.close()
            log(f"[bench] muscl {k}: {out[k]['Mcell_s']:.1f} Mcell/s ({r['step_ms']:.3f} ms/step, stage 2 "
                f"{out[k]['stage2_GBps']:.0f} GB/s)")
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        m = 32
        Uc, dc = muscl_turbulence_state(m, ng)
        O.muscl_step_3d(Uc, dc, ng, 0.2 / m, 1.4, 1.5, 1.0 / m)
        t0 = time.perf_counter()
        k = 0
        while time.perf_counter() - t0 < 5.0:
            O.muscl_step_3d(Uc, dc, ng, 0.2 / m, 1.4, 1.5, 1.0 / m)
            k += 1
        el = (time.perf_counter() - t0) / k
        out["cpu_baseline"] = {"value": m ** 3 / el / 1e6, "unit": "Mcell/s", "cores": 1, "kind": "port",
                               "sample": f"{k} Float64 steps of the oracle at {m}^3 + {ng} ghost zones"}
        log(f"[bench] muscl CPU oracle {out['cpu_baseline']['value']:.2f} Mcell/s")
    return out


def gravity_leg(V, args):
    """Self-gravity on a 3D periodic box (n^3, gamma=5/3, G=1): a Jeans-unstable
    density ripple rho = 1 + 1e-3 (cos 2πx + cos 2πy + cos 2πz) plus 1e-4 random
    cells, p = 0.06, v = 0. Times the device CG Poisson solve from φ = 0 (tol 1e-8)
    and warm-started after a step, and the SSP-RK2 step with the gravity source
    against the same step without gravity, FP64 fast profile."""
    import numpy as np
    n = args.gravity_n
    x = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    rng = np.random.default_rng(5)
    rho = 1.0 + 1e-3 * (np.cos(2 * np.pi * X) + np.cos(2 * np.pi * Y) + np.cos(2 * np.pi * Z)) \
        + 1e-4 * rng.standard_normal(X.shape)
    W = np.stack([rho, 0 * X, 0 * X, 0 * X, 0.06 + 0 * X])
    U = V.prim2cons_array(W, 5.0 / 3.0, np.float64)
    del W, X, Y, Z, rho
    prob = V.Problem("jeans3d", (n,) * 3, ((0.0, 1.0),) * 3, 5.0 / 3.0, "periodic",
                     lambda x, y, z: (1.0, 0.0, 0.0, 0.0, 1.0), 1.0, 0.3)
    sim = V.Simulation(V.UniformMesh((n,) * 3, prob.domain), prob)
    sim.set_mode(args.mode)
    sim.upload(U)
    # the same step without gravity, for the source's cost
    r_plain = sim.time_steps(args.gravity_steps)
    sim.upload(U)
    grav = V.enable_gravity(sim, G=1.0, bcs="periodic", tol=1e-8, maxiter=2000)
    sim.sync()
    t0 = time.perf_counter()
    it_cold, rr_cold = grav.solve()
    cold = time.perf_counter() - t0
    r_grav = sim.time_steps(args.gravity_steps)
    t0 = time.perf_counter()
    it_warm, rr_warm = grav.solve()
    warm = time.perf_counter() - t0
    ncell = n ** 3
    per_it = cold / max(it_cold, 1)
    # algorithmic HBM bytes per CG iteration: matvec 16 B/cell (read p, write Ap),
    # update 48 B/cell (read φ, r, p, Ap; write φ, r), direction 24 B/cell (read r, p; write p)
    bytes_it = 88 * ncell
    res = {"workload": f"3D {n}^3 periodic Jeans ripple, gamma=5/3, G=1, CG tol 1e-8, FP64 {args.mode} profile",
           "cells": ncell,
           "poisson_cold": {"iterations": it_cold, "relres": rr_cold, "seconds": cold,
                            "ms_per_iteration": 1e3 * per_it, "GBps_algorithmic": bytes_it / per_it / 1e9},
           "poisson_warm_after_step": {"iterations": it_warm, "relres": rr_warm, "seconds": warm},
           "step_ms_with_gravity": r_grav["step_ms"], "step_ms_without": r_plain["step_ms"],
           "zone_updates_per_s_with_gravity": ncell / (r_grav["step_ms"] * 1e-3)}
    log(f"[bench] gravity {n}^3: CG {it_cold} it in {cold * 1e3:.1f} ms ({1e3 * per_it:.3f} ms/it, "
        f"{bytes_it / per_it / 1e9:.0f} GB/s algorithmic), warm {it_warm} it; step {r_grav['step_ms']:.3f} ms "
        f"with gravity vs {r_plain['step_ms']:.3f} ms without")
    if not args.no_cpu_baseline:
        if ROOT not in sys.path:
            sys.path.insert(0, ROOT)
        from oracle import oracle as O
        m = 64
        g = O.make_grid((m,) * 3, [(0.0, 1.0)] * 3, "periodic", 5.0 / 3.0)
        xs = (np.arange(m) + 0.5) / m
        Zs, Ys, Xs = np.meshgrid(xs, xs, xs, indexing="ij")
        Us = np.zeros((5, m, m, m))
        # the ripple alone is one eigenvector of the discrete Laplacian (CG would stop after one
        # iteration); the seeded cell noise of the device workload makes it a real solve
        Us[0] = 1.0 + 1e-3 * (np.cos(2 * np.pi * Xs) + np.cos(2 * np.pi * Ys) + np.cos(2 * np.pi * Zs)) \
            + 1e-4 * np.random.default_rng(5).standard_normal(Xs.shape)
        Us[4] = 0.06 / (5.0 / 3.0 - 1)
        phi = np.zeros(m ** 3)
        t0 = time.perf_counter()
        it, rr = O.poisson_solve(g, 7, Us, 1.0, True, 2000, 1e-8, phi)
        el = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": m ** 3 * it / el, "unit": "cell-iterations/s", "cores": 1, "kind": "port",
                               "sample": f"oracle CG solve of a {m}^3 noisy ripple from phi=0 to relres {rr:.1e}, "
                                         f"{it} iterations, {el:.2f} s"}
        res["gpu_cell_iterations_per_s"] = ncell / per_it
        log(f"[bench] gravity CPU {m ** 3 * it / el / 1e6:.1f} M cell-iterations/s vs GPU {ncell / per_it / 1e9:.2f} G")
    grav.close()
    sim.close()
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--tile", type=int, default=-1, help="tile variant (-1: library default)")
    ap.add_argument("--zchunk", type=int, default=0)
    ap.add_argument("--mode", choices=("fast", "exact"), default="fast",
                    help="arithmetic profile of the headline FP64 run (fast: within 1e-12 of the "
                         "reference arithmetic; exact: bit-identical)")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-other-mode", action="store_true",
                    help="skip the FP64 run in the other arithmetic profile")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sod", action="store_true", help="skip the Sod-400 leg")
    ap.add_argument("--no-blocks", action="store_true", help="skip the batched-blocks leg")
    ap.add_argument("--skip-uniform", action="store_true",
                    help="skip the 256^3 legs (profiling runs of the other legs only; no headline value)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint rollout leg")
    ap.add_argument("--no-sedov", action="store_true", help="skip the dynamic-AMR Sedov leg")
    ap.add_argument("--no-gravity", action="store_true", help="skip the self-gravity leg")
    ap.add_argument("--no-muscl", action="store_true", help="skip the PPMKernels MUSCL leg")
    ap.add_argument("--muscl-n", type=int, default=256)
    ap.add_argument("--muscl-steps", type=int, default=20)
    ap.add_argument("--gravity-n", type=int, default=256)
    ap.add_argument("--gravity-steps", type=int, default=20)
    ap.add_argument("--adjoint-n", type=int, default=512)
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--fd-directions", type=int, default=16)
    ap.add_argument("--blocks", type=int, default=3000, help="blocks of the batched-blocks leg")
    ap.add_argument("--block-steps", type=int, default=20)
    ap.add_argument("--no-autotune", action="store_true", help="batched-blocks leg: keep the default launch shape")
    ap.add_argument("--block-methods", default="graph,persistent,per_block",
                    help="launch methods of the blocks leg, in order (Float32 runs the first only)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also time 3 device-to-device copies of 1 GiB (counter calibration)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build measures one GPU only: the flux-divergence path is "
                                   "benchmarked on a single MI355X; multi-GPU sharding is design-only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    V = load_package()
    if V.device_count() < 1:
        raise SystemExit("no HIP device visible")
    from vespa_jl_amd import problems as P

    out = {}
    calib = None
    if args.calib_copy:
        nb = 1 << 30
        a, b = V.DeviceBuffer(nb), V.DeviceBuffer(nb)
        b.copy_from(a)
        t0 = time.perf_counter()
        for _ in range(3):
            b.copy_from(a)
        el = (time.perf_counter() - t0) / 3
        calib = {"copy_bytes": nb, "copy_GBps_read_plus_write": 2 * nb / el / 1e9}
        a.free()
        b.free()
        log(f"[bench] D2D copy 1 GiB: {calib['copy_GBps_read_plus_write']:.0f} GB/s (read+write)")
    if not args.skip_uniform:
        n = args.n
        ncell = n ** 3
        log(f"[bench] building {n}^3 turbulence-like initial state")
        U64 = P.turbulence_state(n)

        log(f"[bench] FP64 run ({args.mode} profile)")
        sim = make_sim(V, n, np.float64, U64, args.tile, args.zchunk, args.mode)
        launch = sim.launch()
        wall, ev = timed_run(V, sim, args.steps, args.warmup)
        value = ncell * args.steps / wall
        ms_per_step = 1e3 * wall / args.steps
        stage2_bytes = 15 * 8 * ncell
        stage1_bytes = 10 * 8 * ncell
        ach = stage2_bytes / (ev["stage2_ms"] * 1e-3) / 1e9
        log(f"[bench] FP64 {value / 1e9:.3f} G zone-updates/s, {ms_per_step:.3f} ms/step, "
            f"stage1 {ev['stage1_ms']:.3f} ms, stage2 {ev['stage2_ms']:.3f} ms, stage2 {ach:.0f} GB/s")
        U64_end = sim.state() if not (args.no_fp32 and args.no_other_mode) else None
        sim.close()

        traffic, traffic_src = None, None
        tfile = os.path.join(ROOT, "profiles", "traffic_rk_stage2_256cubed_f64.json")
        if os.path.exists(tfile) and n == 256 and args.mode == "fast" and args.tile < 0 and args.zchunk == 0:
            try:
                rec = json.load(open(tfile))
                # counter bytes of a rocprofv3 --pmc pass over this kernel; reported only
                # while the kernel sources still hash to what was profiled
                if rec.get("stage_sources_sha256") == stage_sources_hash():
                    traffic = rec.get("hbm_bytes_per_launch")
                    traffic_src = "profiles/traffic_rk_stage2_256cubed_f64.json"
            except (ValueError, OSError):
                traffic = None

        out = {
            "metric": "zone-updates/s & %HBM peak, 3D 256³ PLM+HLLC RK2; adjoint/forward time ratio",
            "value": value,
            "unit": "zone-updates/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": ms_per_step,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f64",
            "arithmetic_profile": args.mode,
            "data": "synthetic",
            "config": {"workload": f"3D {n}^3 periodic smooth turbulence-like state (rho=1+0.1*delta, "
                                   f"|k|<=8 random-phase modes, v_rms=0.1 c_s, p=1), gamma=5/3, "
                                   f"PLM-minmod+HLLC SSP-RK2, CFL 0.3 with device dt every step, FP64",
                       "grid": [n, n, n], "fields": "5 conserved, SoA", "tile": launch["tile"],
                       "zchunk": launch["zchunk"], "workgroups_per_stage": launch["workgroups"]},
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                         "traffic_source": traffic_src,
                         "kernel": "rk_stage_kernel<double,3,...,STAGE2=true>",
                         "bytes_per_launch": stage2_bytes, "mean_launch_ms": ev["stage2_ms"]},
            "step_detail": {"stage1_ms": ev["stage1_ms"], "stage2_ms": ev["stage2_ms"],
                            "event_step_ms": ev["step_ms"],
                            "stage1_GBps": stage1_bytes / (ev["stage1_ms"] * 1e-3) / 1e9,
                            "step_algorithmic_GBps": 200 * ncell / (ms_per_step * 1e-3) / 1e9},
        }

        if calib is not None:
            out["calibration"] = calib

        if not args.no_other_mode:
            other = "exact" if args.mode == "fast" else "fast"
            log(f"[bench] FP64 run ({other} profile)")
            so = make_sim(V, n, np.float64, U64, args.tile, args.zchunk, other)
            wo, evo = timed_run(V, so, args.steps, args.warmup)
            Uo_end = so.state()
            so.close()
            vo = ncell * args.steps / wo
            rel = (np.abs(Uo_end - U64_end).reshape(5, -1).max(axis=1) /
                   np.abs(U64_end).reshape(5, -1).max(axis=1)) if U64_end is not None else None
            out["fp64_" + other] = {"value": vo, "unit": "zone-updates/s", "ms_per_step": 1e3 * wo / args.steps,
                                    "stage1_ms": evo["stage1_ms"], "stage2_ms": evo["stage2_ms"],
                                    "stage2_GBps": stage2_bytes / (evo["stage2_ms"] * 1e-3) / 1e9,
                                    "max_rel_diff_vs_headline_per_field":
                                        rel.tolist() if rel is not None else None}
            log(f"[bench] FP64 {other}: {vo / 1e9:.3f} G zone-updates/s")

        if not args.no_fp32:
            log("[bench] FP32 run")
            U32 = U64.astype(np.float32)
            s32 = make_sim(V, n, np.float32, U32, args.tile, args.zchunk, args.mode)
            wall32, ev32 = timed_run(V, s32, args.steps, args.warmup)
            v32 = ncell * args.steps / wall32
            U32_end = s32.state().astype(np.float64)
            s32.close()
            den = np.abs(U64_end).mean(axis=(1, 2, 3))
            l1 = (np.abs(U32_end - U64_end).mean(axis=(1, 2, 3)) / den).tolist()
            out["fp32"] = {"arithmetic_profile": args.mode, "value": v32, "unit": "zone-updates/s", "ratio_vs_fp64": v32 / value,
                           "ms_per_step": 1e3 * wall32 / args.steps, "stage1_ms": ev32["stage1_ms"],
                           "stage2_ms": ev32["stage2_ms"],
                           "stage2_GBps": 15 * 4 * ncell / (ev32["stage2_ms"] * 1e-3) / 1e9,
                           "rel_L1_vs_fp64_per_field": dict(zip(
                               ["density", "momentum_x", "momentum_y", "momentum_z", "energy"], l1)),
                           "after_steps": args.warmup + args.steps}
            log(f"[bench] FP32 {v32 / 1e9:.3f} G zone-updates/s (x{v32 / value:.2f}), rel L1 {l1}")

        if not args.no_cpu_baseline:
            log("[bench] CPU baseline (oracle, OpenMP)")
            out["cpu_baseline"] = cpu_baseline(U64, n)
            log(f"[bench] CPU {out['cpu_baseline']['value'] / 1e6:.2f} M zone-updates/s")
        else:
            out["cpu_baseline"] = None

    if not args.no_sod:
        out["sod"] = sod_leg(V, args)
    if not args.no_blocks:
        out["batched_blocks"] = blocks_leg(V, args)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg(V, args)
    if not args.no_sedov:
        out["sedov_amr"] = sedov_leg(V, args)
    if not args.no_gravity:
        out["self_gravity"] = gravity_leg(V, args)
    if not args.no_muscl:
        out["muscl_rk2"] = muscl_leg(V, args)

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
