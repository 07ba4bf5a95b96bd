"""Benchmark: SPH particle-steps/s of the coupled WCSPH-DEM step on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2]

(--workload e0|e1 measures the Eulerian lid-driven cavity instead: point-
steps/s of EISPH_OneStep, see bench_cavity.)

Workload (default c2): 1,562,500 fluid particles (dx = 0.004), 1,080,816
dummy-wall particles, 2,000 DEM spheres, 14,688 DEM-wall spheres,
dt = 2.5e-5 s, synthetic initial state from the dam-break generator.  A
step is one pass of the hot path -- grid rebuilds plus SPHDEM_OneStep --
over the resident state; `value` = fluid particles x K / wall seconds of K
steps (inputs already in HBM).

Also reported:
  roofline      the pass with the largest GPU time in the timed region:
                its algorithmic bytes per launch (each array it reads or
                writes counted once, per the table below) over its average
                launch time from HIP events recorded around every launch in
                the timed region, against 8 TB/s.
  cpu_baseline  the CPU oracle (C, OpenMP on this host's cores) timed on one
                coupled step of the same workload.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import torch

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# FP32 operations of one porosity term (KERNEL_SPHDEM_interaction.py:80-83
# with Kernel_KNL.py:17-24), the per-subject constant 21/(16 pi h^3) taken
# as one multiply: the difference and its length (3 sub, 3 mul, 2 add,
# sqrt), q = r/h (1), u = 1 - q/2 (2), u^4 (2), 1 + 2q (2), C u^4 (1 + 2q)
# (2), m/rho (1), V W (1) and the sum (1)
POROSITY_FLOP_PER_PAIR = 21

# Algorithmic bytes per launch of each pass: (per fluid, per wall, per DEM).
PASS_BYTES = {
    "grid_build": (68.0, 0.0, 0.0),
    "fluid_density": (56.0, 16.0, 0.0),
    "wall_density": (16.0, 60.0, 0.0),
    "fluid_force": (92.0, 40.0, 0.0),
    "porosity": (28.0, 0.0, 32.0),
    "interaction_dem": (52.0, 0.0, 120.0),
    "fluid_final": (88.0, 0.0, 44.0),
    "dem_csr": (0.0, 0.0, 40.0),
    "dem_contact": (0.0, 0.0, 160.0),
    "step_dem": (0.0, 0.0, 100.0),
    "other": (0.0, 0.0, 0.0),
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: workload's)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c2", choices=("c0", "c1", "c2", "c3", "e0", "e1"))
    ap.add_argument("--graph", action="store_true",
                    help="replay captured hipGraphs (always on for the c3 ensemble)")
    ap.add_argument("--instances", type=int, default=64, help="c3: number of instances")
    ap.add_argument("--tune-steps", type=int, default=3,
                    help="steps per autotuner candidate during warm-up (0: library defaults)")
    ap.add_argument("--surrogate-sim-time", type=float, default=0.5,
                    help="simulated seconds of the C1 surrogate-vs-plain rollout (0: skip)")
    ap.add_argument("--math", choices=("exact", "fast"), default="fast",
                    help="W/dW arithmetic of the neighbour sums (parity-tested both ways)")
    ap.add_argument("--cpu-steps", type=int, default=1, help="oracle steps for cpu_baseline")
    ap.add_argument("--cpu-stride", type=int, default=None,
                    help="oracle subject stride (default 8 for c2, else 1)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph-leg", action="store_true",
                    help="skip the hipGraph replay timing after the eager timed region")
    ap.add_argument("--traffic", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                       "profiles", "porosity_c2_traffic.json"),
                    help="JSON file with measured HBM bytes per launch of the dominant pass "
                         "(from rocprofv3 FETCH_SIZE/WRITE_SIZE, scaled by a copy of known size)")
    ap.add_argument("--counters", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                        "profiles", "porosity_c2_valu.json"),
                    help="JSON file with rocprofv3 instruction counters of the dominant pass "
                         "(matrix products per launch, VALU issue fraction)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only copy a buffer of this many MiB 5 times (counter calibration)")
    args = ap.parse_args(argv)

    if args.calibrate_copy_mib:
        n = args.calibrate_copy_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes": n * 4}))
        return 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        return 2

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from warp_sph_gpt_amd import _lib
    from warp_sph_gpt_amd.engine import CoupledEngine
    from warp_sph_gpt_amd.workloads import WORKLOADS

    if args.workload == "c3":
        return bench_ensemble(args)
    if args.workload in ("e0", "e1"):
        return bench_cavity(args)

    dev = torch.device("cuda:0")
    t0 = time.time()
    wl = WORKLOADS[args.workload](device="cuda:0")
    P, B, D, W = (s.to(dev) for s in wl.structs)
    steps = args.steps if args.steps is not None else wl.steps
    log(f"[bench] {wl.name}: {P.n} fluid, {B.n} wall, {D.n} DEM, {W.n} DEM-wall; {wl.note}; "
        f"setup {time.time() - t0:.1f}s")
    eng = CoupledEngine(wl.solv, P, B, D, W, fast_math=args.math == "fast")
    tuned = None
    if args.tune_steps > 0:
        from warp_sph_gpt_amd import autotune
        tuned = autotune.autotune(eng.step, steps=args.tune_steps, log=log)["tuning"]
    for _ in range(args.warmup):
        eng.step()
    torch.cuda.synchronize()
    if args.graph:
        eng.capture()
    log(f"[bench] warmup {args.warmup} steps done")

    torch.cuda.synchronize()
    with _lib.PassTimer() as timer:
        start = time.perf_counter()
        eng.run(steps)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - start
    eng.check()
    diag = eng.diagnostics()
    log(f"[bench] {steps} steps in {elapsed:.3f}s; per-pass ms: "
        + ", ".join(f"{k}={v:.1f}" for k, v in timer.ms.items() if v > 0))
    # the same number of steps again as hipGraph replays (no per-pass events)
    graph_run = None
    if not args.graph and not args.no_graph_leg:
        eng.capture()
        eng.run(2)
        torch.cuda.synchronize()
        start = time.perf_counter()
        eng.run(steps)
        torch.cuda.synchronize()
        g_el = time.perf_counter() - start
        eng.check()
        graph_run = {"ms_per_step": round(g_el / steps * 1e3, 4),
                     "value": round(P.n * steps / g_el, 1), "steps": steps,
                     "note": "same engine, one hipGraph replay per step after the eager "
                             "timed region; value above is the eager run's"}
        log(f"[bench] {steps} graph steps in {g_el:.3f}s")

    value = P.n * steps / elapsed
    dom = max(timer.ms, key=lambda k: timer.ms[k])
    count = max(timer.count[dom], 1)
    avg_s = timer.ms[dom] / count * 1e-3
    bf, bb, bd = PASS_BYTES[dom]
    bytes_launch = bf * P.n + bb * B.n + bd * D.n
    achieved = bytes_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    here = os.path.dirname(os.path.abspath(__file__))
    traffic = traffic_src = None
    if args.traffic and os.path.exists(args.traffic):
        with open(args.traffic) as f:
            t = json.load(f)
        if t.get("kernel") == dom and t.get("workload") == wl.name:
            traffic = int(t["bytes_per_launch"])
            traffic_src = os.path.relpath(args.traffic, here)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                "traffic_source": traffic_src,
                "kernel": dom, "avg_launch_ms": round(avg_s * 1e3, 4),
                "bytes_per_launch": int(bytes_launch)}
    # pair rate and FP32 rate of the dominant pass: pairs per launch from its
    # counter profile (16 x 16 pairs per matrix product) over this run's
    # launch time, at the reference term's flops per pair
    counters = None
    if args.counters and os.path.exists(args.counters) and avg_s > 0:
        with open(args.counters) as f:
            c = json.load(f)
        if c.get("kernel") == dom and c.get("workload") == wl.name and c.get("math") == args.math:
            src = os.path.relpath(args.counters, here)
            pairs = c["pairs_evaluated_per_launch"]
            counters = {"pairs_evaluated_per_launch": pairs,
                        "pair_rate": round(pairs / avg_s, 1),
                        "valu_busy_frac": c["valu_busy_frac"], "source": src}
            # the pass is bound by the vector ALUs, not by HBM bytes: its
            # FP32 rate against the vector FP32 peak, and how busy the VALUs
            # were in the counter run
            tflops = pairs * POROSITY_FLOP_PER_PAIR / avg_s / 1e12
            roofline["limiter"] = {"what": "VALU busy (4 SQ_ACTIVE_INST_VALU / (1024 SIMDs x "
                                           "GRBM_GUI_ACTIVE per XCD))",
                                   "frac": c["valu_busy_frac"], "source": src}
            roofline["valu"] = {"achieved": round(tflops, 2), "peak": FP32_PEAK_TFLOPS,
                                "unit": "TFLOP/s", "frac": round(tflops / FP32_PEAK_TFLOPS, 4),
                                "flop_per_pair": POROSITY_FLOP_PER_PAIR,
                                "pairs_per_launch": pairs, "busy_frac": c["valu_busy_frac"],
                                "source": src}

    cpu = cpu1 = None
    if not args.no_cpu_baseline:
        stride = args.cpu_stride or (8 if wl.name == "c2" else 1)
        cpu = cpu_baseline(wl, args.cpu_steps, stride)
        from oracle import oracle as O
        threads = O.num_threads()
        O.set_num_threads(1)
        try:
            cpu1 = cpu_baseline(wl, args.cpu_steps, 64 if wl.name == "c2" else stride)
        finally:
            O.set_num_threads(threads)
    sur = None
    if args.surrogate_sim_time > 0:
        del eng
        torch.cuda.empty_cache()
        sur = surrogate_rollout(args.surrogate_sim_time)

    line = {
        "metric": "SPH particle-steps/s at 1.56M ptl; surrogate speedup at <=1% rho error",
        "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(elapsed / steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{wl.name}: coupled WCSPH-DEM dam break", "fluid": P.n,
                   "wall": B.n, "dem": D.n, "dem_wall": W.n, "dt": wl.solv.dt,
                   "graph": bool(args.graph), "math": args.math, "tuning": tuned,
                   "parallelism": "single GPU"},
        "roofline": roofline, "cpu_baseline": cpu, "cpu_baseline_1thread": cpu1,
        "graph": graph_run, "surrogate": sur,
        "dominant_pass_counters": counters,
        "passes_ms_per_step": {k: round(v / steps, 4) for k, v in timer.ms.items() if v > 0},
        # these run on other streams beside the SPH passes and the porosity:
        # their event intervals include the work they overlap
        "passes_overlapped": ["dem_csr", "dem_contact", "interaction_dem"] + (
            ["wall_density"] if _lib.get_tuning()["wall_beside"] else []) + (
            ["fluid_force"] if _lib.get_tuning()["force_beside"] else []),
        "diagnostics_fp64": diag,
    }
    print(json.dumps(line), flush=True)
    return 0


def roofline_of(timer, steps_counted, nf, nb, nd, traffic=None) -> dict:
    dom = max(timer.ms, key=lambda k: timer.ms[k])
    count = max(timer.count[dom], 1)
    avg_s = timer.ms[dom] / count * 1e-3
    bf, bb, bd = PASS_BYTES[dom]
    bytes_launch = bf * nf + bb * nb + bd * nd
    achieved = bytes_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
            "kernel": dom, "avg_launch_ms": round(avg_s * 1e3, 4),
            "bytes_per_launch": int(bytes_launch)}


def bench_ensemble(args) -> int:
    """C3: the batched ensemble, autotuned launch shapes, one hipGraph replay
    per step for all instances.  value = fluid particles of all instances x
    steps / wall seconds."""
    from warp_sph_gpt_amd import _lib, autotune
    from warp_sph_gpt_amd.ensemble import EnsembleEngine
    from warp_sph_gpt_amd.workloads import ENSEMBLES

    t0 = time.time()
    wl = ENSEMBLES["c3"](device="cuda:0", n_inst=args.instances)
    ens = EnsembleEngine(wl.solv, wl.scenes, device="cuda:0", fast_math=args.math == "fast")
    nf, nb = ens.P.n, ens.B.n
    nd = ens.D.n if ens.D is not None else 0
    steps = args.steps if args.steps is not None else wl.steps
    log(f"[bench] c3: {wl.note}; {nb} wall, {ens.W.n if ens.W is not None else 0} DEM-wall; "
        f"setup {time.time() - t0:.1f}s")
    tuned = None
    if args.tune_steps > 0:
        tuned = autotune.autotune(ens.step, steps=args.tune_steps, log=log)["tuning"]
    for _ in range(args.warmup):
        ens.step()
    torch.cuda.synchronize()
    with _lib.PassTimer() as timer:          # eager per-pass breakdown (2 steps)
        for _ in range(2):
            ens.step()
        torch.cuda.synchronize()
    ens.capture()
    ens.step()
    torch.cuda.synchronize()
    log(f"[bench] autotuned, warmed up and captured at step {ens.step_index}")
    start = time.perf_counter()
    ens.run(steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - start
    ens.check()
    log(f"[bench] {steps} graph steps in {elapsed:.3f}s")
    line = {
        "metric": "SPH particle-steps/s, batched ensemble (all instances)",
        "value": round(nf * steps / elapsed, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": round(elapsed / steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"c3: {wl.note}", "instances": ens.n, "fluid": nf, "wall": nb,
                   "dem": nd, "dt": wl.solv.dt, "graph": True, "math": args.math,
                   "tuning": tuned, "parallelism": "single GPU"},
        "roofline": roofline_of(timer, 2, nf, nb, nd),
        "cpu_baseline": None,
        "passes_ms_per_step_eager": {k: round(v / 2, 4) for k, v in timer.ms.items() if v > 0},
    }
    print(json.dumps(line), flush=True)
    return 0


def surrogate_rollout(sim_time: float) -> dict:
    """C1 (100,000 fluid, 200 DEM): the plain GPU path and the surrogate
    rollout (judge and fallback every 50 calls, 1 % density) over the same
    simulated time; simulated seconds per wall second of each and their ratio."""
    from warp_sph_gpt_amd.surrogate import MPNN, WEIGHTS, load_params, rollout_c1
    if not os.path.exists(WEIGHTS):
        log("[bench] surrogate weights missing; skipped")
        return None
    net = MPNN(params=load_params(WEIGHTS))
    r = rollout_c1(net, sim_time=sim_time, judge_every=50, density_tol=0.01)
    keep = ("sim_time", "calls", "speedup", "plain_sim_per_wall", "surrogate_sim_per_wall",
            "surrogate_calls", "plain_calls", "judged", "rejected", "unhealthy",
            "velocity_err_mean", "velocity_err_max", "density_err_mean", "density_err_max")
    out = {k: (round(r[k], 6) if isinstance(r[k], float) and r[k] == r[k]
               and abs(r[k]) != float("inf") else (None if isinstance(r[k], float) else r[k]))
           for k in keep}
    out["workload"] = "c1: 100,000 fluid, 200 DEM; judge every 50 calls at 1% density"
    out["final_ke_rel"] = abs(r["final_surrogate"]["fluid_ke"] - r["final_plain"]["fluid_ke"]) \
        / r["final_plain"]["fluid_ke"]
    out["final_front_rel"] = abs(r["final_surrogate"]["front"] - r["final_plain"]["front"]) \
        / r["final_plain"]["front"]
    de, ve = out["density_err_max"], out["velocity_err_max"]
    out["bar"] = {"speedup": 5.0, "density_err": 0.01, "velocity_err": 0.02,
                  "met": bool(out["speedup"] >= 5.0 and de is not None and de <= 0.01
                              and ve is not None and ve <= 0.02)}
    # what a call costs (stages of an eager call, a replayed call, ten plain
    # substeps replayed) and the network kernels' matrix-core use from their
    # counter profile
    out["call_costs"] = r.get("call_costs")
    mf = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                      "surrogate_c1_mfma.json")
    if os.path.exists(mf):
        with open(mf) as f:
            out["network_mfma"] = dict(json.load(f), source="profiles/surrogate_c1_mfma.json")
    log(f"[bench] surrogate: {json.dumps(out)}")
    return out


# Algorithmic bytes per launch of the cavity passes: (per fluid point, per
# neighbour-list entry, per ghost).  A point's own fields (row length and
# slice offset included) and each list entry (id 4 B + (gx, gz, A_ij, V_j)
# 16 B) are counted once; neighbour gathers re-read arrays already counted.
EISPH_PASS_BYTES = {
    "eisph_advec": (4 + 12 + 24, 20, 16),          # offset, vel | acc, vel_star; vel_bc, mirror
    "eisph_ppe": (4 + 12 + 4 + 16, 20, 28),         # offset, vel_star, pres | div, A, b, Ap
    "eisph_solve": (16 + 8, 0, 0),                  # b, Ap, A, p | b, p
    "eisph_correct": (4 + 4 + 12 + 4 + 12 + 24, 20, 8),  # offset, p, vel*, rho, vel | vel, acc
    "eisph_finish": (0, 0, 28),                     # ghosts: vel_bc, mirror | vel
    # one workgroup, per step: the three list passes (fields stay in LDS)
    "eisph_steps": (0, 3 * 20, 0),
}


def bench_cavity(args) -> int:
    """The Eulerian lid-driven cavity (source/EISPH.py): value = fluid points
    x K / wall seconds of K steps (diagnostics off, as between output steps
    of run_eisph), fixed grids, KGC and neighbour list built before timing."""
    from warp_sph_gpt_amd import _lib
    from warp_sph_gpt_amd import eisph as G
    from warp_sph_gpt_amd.workloads import CAVITIES, cavity_case

    case = CAVITIES[args.workload]
    t0 = time.time()
    gen, solv, P, B = cavity_case(case.name)
    dims = None if case.cells <= 64 else (512, 1, 512)
    gs, gb = G.prepare(solv, P, B, dims=dims)
    L = G.neighbours(P, B, gs, gb, solv.support, solv.h)
    steps = args.steps if args.steps is not None else case.steps
    log(f"[bench] {case.name}: {P.n} fluid, {B.n} ghost, {L.entries} list entries; "
        f"setup {time.time() - t0:.1f}s")
    for _ in range(args.warmup):
        G.EISPH_OneStep(solv, P, B, gs, gb, diagnostics=False)
    # a cloud that fits one workgroup runs 100 steps (the shipped output
    # interval) per launch of the multi-step kernel, unless --graph asks for
    # replays of the five-launch step
    per_call = min(steps, 100) if G.one_group(P, B) and not args.graph else 0
    if per_call:
        steps = -(-steps // per_call) * per_call
        G.EISPH_Steps(solv, P, B, gs, gb, 1, diagnostics=False)
    torch.cuda.synchronize()
    per_graph = 0
    if args.graph:
        per_graph = min(steps, 100)
        steps = -(-steps // per_graph) * per_graph
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with torch.cuda.graph(graph, stream=side):
                for _ in range(per_graph):
                    G.EISPH_OneStep(solv, P, B, gs, gb, diagnostics=False)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
    log(f"[bench] warmup {args.warmup} steps done")

    torch.cuda.synchronize()
    if per_call:
        with _lib.PassTimer() as timer:
            start = time.perf_counter()
            for _ in range(steps // per_call):
                G.EISPH_Steps(solv, P, B, gs, gb, per_call, diagnostics=False)
            torch.cuda.synchronize()
            elapsed = time.perf_counter() - start
        timed_from = "HIP events around every launch of the timed region"
    elif per_graph:
        start = time.perf_counter()
        for _ in range(steps // per_graph):
            graph.replay()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - start
        # per-pass launch times from profiled eager steps after the timed region
        with _lib.PassTimer() as timer:
            for _ in range(min(steps, 50)):
                G.EISPH_OneStep(solv, P, B, gs, gb, diagnostics=False)
            torch.cuda.synchronize()
        timed_from = "profiled eager steps after the timed graph replays"
    else:
        with _lib.PassTimer() as timer:
            start = time.perf_counter()
            for _ in range(steps):
                G.EISPH_OneStep(solv, P, B, gs, gb, diagnostics=False)
            torch.cuda.synchronize()
            elapsed = time.perf_counter() - start
        timed_from = "HIP events around every launch of the timed region"
    G.EISPH_OneStep(solv, P, B, gs, gb, diagnostics=True)
    torch.cuda.synchronize()
    finite = all(bool(torch.isfinite(getattr(P, k)).all()) for k in ("vel", "pres"))
    passes = {k: v for k, v in timer.ms.items() if k.startswith("eisph_")}
    log(f"[bench] {steps} steps in {elapsed:.3f}s; per-pass ms: "
        + ", ".join(f"{k}={v:.2f}" for k, v in passes.items() if v > 0))

    dom = max(passes, key=lambda k: passes[k])
    count = max(timer.count[dom], 1)
    avg_s = passes[dom] / count * 1e-3
    bf, be, bg = EISPH_PASS_BYTES[dom]
    bytes_launch = (bf * P.n + be * L.entries + bg * B.n) * (per_call if dom == "eisph_steps" else 1)
    achieved = bytes_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    traffic = None
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                         f"eisph_{case.name}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        if t.get("kernel") == dom:
            traffic = int(t["bytes_per_launch"])
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                "kernel": dom, "avg_launch_ms": round(avg_s * 1e3, 4),
                "bytes_per_launch": int(bytes_launch), "launch_times": timed_from}
    cpu = None if args.no_cpu_baseline else cpu_baseline_cavity(case, args.cpu_steps)
    line = {
        "metric": "EISPH cavity point-steps/s", "value": round(P.n * steps / elapsed, 1),
        "unit": "point-steps/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": round(elapsed / steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{case.name}: {case.note}", "fluid": P.n, "ghost": B.n,
                   "list_entries": L.entries, "dt": solv.dt, "graph_steps": per_graph,
                   "steps_per_launch": per_call,
                   "parallelism": "single GPU"},
        "roofline": roofline, "cpu_baseline": cpu, "finite": finite,
        "passes_ms_per_launch": {k: round(v / max(timer.count[k], 1), 4)
                                 for k, v in passes.items() if v > 0},
    }
    print(json.dumps(line), flush=True)
    return 0


def cpu_baseline_cavity(case, n_steps: int) -> dict:
    """The C oracle (OpenMP) on the same cavity: n_steps steps without
    diagnostics after its fixed grids and KGC are built."""
    from oracle import eisph as E
    from oracle import oracle as O
    from warp_sph_gpt_amd.workloads import cavity_case

    gen, solv, P, B = cavity_case(case.name, device="cpu")
    oP, oB = E.from_struct(P), E.from_struct(B)
    gs, gb = E.build_grids(solv, oP, oB, dims=None if case.cells <= 64 else (512, 1, 512))
    n_steps = max(1, n_steps if case.cells > 64 else 100 * n_steps)
    t = time.perf_counter()
    for _ in range(n_steps):
        E.EISPH_OneStep(solv, oP, oB, gs, gb, diagnostics=False)
    el = time.perf_counter() - t
    log(f"[bench] cpu oracle: {n_steps} cavity step(s) in {el:.2f}s on {O.num_threads()} threads")
    return {"value": round(oP.n * n_steps / el, 1), "unit": "point-steps/s",
            "cores": O.num_threads(), "kind": "port",
            "sample": f"{n_steps} step(s) of {case.name} ({oP.n} fluid points), no diagnostics"}


def cpu_baseline(wl, n_steps: int, stride: int) -> dict:
    """The C oracle (OpenMP) on the same workload.  With stride k > 1 the
    fluid/wall neighbour loops visit every k-th subject only; the per-step
    time is then estimated as grid builds + k x (the rest of the step)."""
    from oracle import oracle as O

    structs = [O.OStruct.from_struct(s) for s in wl.structs]
    run = O.Run(wl.solv, *structs)
    O.set_stride(stride)
    t_grid = t_rest = 0.0
    try:
        for _ in range(n_steps):
            t = time.perf_counter()
            run.gs.build(run.P.pos, run.solv.support)
            run.gd.build(run.D.pos, max(2.0 * run.scales[0], 2.0 * run.scales[2]))
            t_grid += time.perf_counter() - t
            t = time.perf_counter()
            O.SPHDEM_OneStep(run.solv, run.P, run.B, run.D, run.W, run.gs, run.gb, run.gd,
                             run.gw, run.step_index, run.scales)
            run.step_index += 1
            t_rest += time.perf_counter() - t
    finally:
        O.set_stride(1)
    est = t_grid + stride * t_rest
    nf = structs[0].n
    log(f"[bench] cpu oracle: {n_steps} step(s): grids {t_grid:.2f}s, rest {t_rest:.2f}s at "
        f"stride {stride}, estimated {est:.2f}s on {O.num_threads()} threads")
    sample = f"{n_steps} coupled step(s) of {wl.name} ({nf} fluid) incl. grid builds"
    if stride > 1:
        sample += (f"; neighbour loops sampled on every {stride}th subject, "
                   f"time = grids + {stride} x rest")
    return {"value": round(nf * n_steps / est, 1), "unit": "particle-steps/s",
            "cores": O.num_threads(), "kind": "port", "sample": sample}


if __name__ == "__main__":
    sys.exit(main())
