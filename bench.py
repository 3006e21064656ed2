#!/usr/bin/env python3
"""Benchmark of the MPS time step (Computer::ForwardTime) on one MI355X.

Workload (default): the 1M-particle dam break, a 2.0 x 1.2 m tank with a
0.5 x 1.0 m water column at l0 = 0.7 mm (1,020,306 fluid + walls), positions
jittered +/-0.02 l0 from a fixed seed, r_e = 2.1 l0, Laplacian radius 3.1 l0,
dt = 1e-4 s, pressure solved to |b - A p| <= 1e-8 |b| (FP64 residual).

A step is one full ForwardTime: neighbour search, two density passes, HS/ECS,
viscosity + gravity, PPE assembly, the pressure solve, pressure gradient and
DS.  The headline solver runs the timed steps; the plain FP64 CG (the
reference algorithm) runs the same steps in a second context for the solve
time comparison.  Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s @1M; solve ms/step learned+mixed vs plain FP64 @res 1e-8"
HBM_PEAK_GBS = 8000.0
REPLAYS = 50
# HBM bytes per launch of the dambreak_1m SpMVs measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE (one run each; FETCH_SIZE scaled by 2, WRITE_SIZE
# by 1: what the counters read for a copy kernel of known size), recorded with
# their commands in profiles/dambreak_1m_step_kernels.md
TRAFFIC = {"cg_spmv32": 116.3e6, "cg_spmv": 389.0e6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_computer(cfg, solver: str, tol: float, weights: str | None = None):
    from openmps_amd import Computer, make_environment, solvers
    c = Computer(cfg["eps"], make_environment(cfg["env"]))
    solvers.configure(c, solver, tol, weights)
    c.add_particles(cfg["particles"])
    return c


def run(c, steps: int, label: str):
    """Runs `steps` ForwardTime calls; returns wall seconds and per-step stats.
    Every call is synchronous at return (stream synchronised inside)."""
    stats = []
    t0 = time.perf_counter()
    for k in range(steps):
        c.forward_time(c._bench_dt)
        s = c.solve_stats()
        stats.append(s)
        log(f"  [{label}] step {k + 1}/{steps}: CG {s['iterations']} it "
            f"({s['outer_iterations']} sweeps), solve {s['solve_ms']:.1f} ms, "
            f"|b-Ax|/|b| {s['rel_residual']:.2e} (from {s['initial_rel_residual']:.2e})")
    return time.perf_counter() - t0, stats


def kernel_table(c):
    out = {}
    for name, s in c.kernel_stats().items():
        if s["launches"] == 0 or s["ms"] <= 0:
            continue
        out[name] = {"launches": s["launches"], "ms_per_launch": s["ms"] / s["launches"],
                     "GBps": s["bytes"] / s["ms"] / 1e6,
                     "bytes_per_launch": s["bytes"] / s["launches"]}
    return out


def profiled_table(c, steps: int, label: str):
    """Per-kernel HIP-event timings over `steps` further steps."""
    c.reset_kernel_stats()
    c.set_profiling(True)
    run(c, steps, f"{label} profiled")
    c.set_profiling(False)
    return kernel_table(c)


def cpu_baseline(cfg, tol: float, budget_s: float):
    """The CPU oracle (C++/OpenMP FP64 restatement of the reference path) on
    one step of the same scene with the same stop rule."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = oracle.Oracle(cfg["env"], cfg["eps"])
    o.set_stop_rule(1, tol)
    o.set_verbose(True)
    o.add_particles(cfg["particles"])
    n = o.size()
    t0 = time.perf_counter()
    o.forward_time(cfg["dt"])
    el = time.perf_counter() - t0
    it = o.cg_stats()["iterations"]
    log(f"  [cpu oracle] 1 step in {el:.2f} s ({it} CG iterations, {threads} threads)")
    return {"value": n / el, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"1 ForwardTime step of the same {n}-particle scene, same stop rule "
                      f"({it} CG iterations), {el:.1f} s", "ms_per_step": el * 1e3}


def run_batch(args) -> int:
    """A batch of 64 uneven independent scenes (scenes.batch): one context
    per scene, each with its own HIP stream; `threads` host threads each step
    a size-balanced share of the contexts (ctypes releases the GIL inside the
    library, so the streams run concurrently).  value = total particle-steps
    per second over the K timed steps of every scene."""
    from concurrent.futures import ThreadPoolExecutor
    from openmps_amd import scenes, solvers
    cfgs = scenes.batch()
    solver = args.solver or "mixed_ml"
    log(f"building {len(cfgs)} scenes ...")
    comps = [make_computer(cfg, solver, args.tol) for cfg in cfgs]
    sizes = [len(cfg["particles"]["x"]) for cfg in cfgs]
    threads = args.threads
    groups = [[] for _ in range(threads)]
    load = [0] * threads
    for k in sorted(range(len(comps)), key=lambda k: -sizes[k]):  # largest first onto the lightest share
        w = load.index(min(load))
        groups[w].append(k)
        load[w] += sizes[k]

    def work(idx, steps):
        for _ in range(steps):
            for k in idx:
                comps[k].forward_time()
        return len(idx)

    with ThreadPoolExecutor(max_workers=threads) as ex:
        if args.warmup:
            list(ex.map(lambda g: work(g, args.warmup), groups))
        log(f"  {len(comps)} scenes, {sum(sizes)} particles, {threads} threads: timing {args.steps} steps")
        t0 = time.perf_counter()
        list(ex.map(lambda g: work(g, args.steps), groups))
        elapsed = time.perf_counter() - t0
    iters = [c.solve_stats()["iterations"] for c in comps]
    total = sum(sizes)
    out = {"metric": "particle-steps/s, 64 uneven dam-break scenes batched on one GPU", "value": total * args.steps / elapsed,
           "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64+f32" if solvers.is_mixed(solver) else "f64", "data": "synthetic",
           "config": {"workload": "batch64", "scenes": len(comps), "particles": total,
                      "particles_min_max": [min(sizes), max(sizes)], "solver": solver, "host_threads": threads,
                      "stop_rule": f"|b-Ax|/|b| <= {args.tol:g}", "dt": "CFL rule, at most 1e-4 s"},
           "last_step_cg_iterations_mean": sum(iters) / len(iters)}
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="dambreak_1m")
    from openmps_amd import solvers
    ap.add_argument("--solver", default=None, choices=sorted(solvers.PRESETS),
                    help="headline solver (default: mixed_ml, the FP32 multilevel-preconditioned "
                         "CG inside FP64 refinement)")
    ap.add_argument("--weights", default=None, help="warm-start MLP weights (.npz)")
    ap.add_argument("--ablation", action="store_true",
                    help="also time zero / previous / learned guesses for both solvers")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-compare", action="store_true", help="skip the plain FP64 comparison")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=60.0)
    ap.add_argument("--threads", type=int, default=8, help="host threads of the batch64 workload")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        return 2

    if args.workload == "batch64":
        return run_batch(args)
    from openmps_amd import scenes
    from openmps_amd import learned as _learned
    have_weights = os.path.exists(args.weights or _learned.DEFAULT_WEIGHTS)
    if args.solver is None:
        args.solver = "mixed_ml"
    log(f"building scene {args.workload} ...")
    cfg = scenes.config(args.workload)
    N = len(cfg["particles"]["x"])
    nfluid = int((cfg["particles"]["type"] == 0).sum())
    log(f"  {N} particles ({nfluid} fluid)")

    main_c = make_computer(cfg, args.solver, args.tol, args.weights)
    main_c._bench_dt = cfg["dt"]
    if args.warmup:
        run(main_c, args.warmup, f"{args.solver} warmup")
    # timed steps run without per-kernel events; the kernel table comes from
    # the same number of further steps with events on
    elapsed, st = run(main_c, args.steps, f"{args.solver} timed")
    kt = profiled_table(main_c, args.steps, args.solver)
    # dominant kernel: the SpMV of the solver that ran the timed steps
    dom = {"none": "cg_iter32", "jacobi": "cg_iter32", "multilevel": "cg_spmv32"}[
        solvers.preconditioner(args.solver)] if solvers.is_mixed(args.solver) else "cg_spmv"
    replay = {}
    if dom in ("cg_spmv32", "cg_spmv"):
        replay[dom] = main_c.replay_spmv(dom, REPLAYS)
        log(f"  {dom}: {replay[dom] * 1e3:.1f} us per launch ({REPLAYS} back-to-back launches)")
    solve_ms = sum(s["solve_ms"] for s in st) / len(st)
    iters = sum(s["iterations"] for s in st) / len(st)

    def summary(stats):
        return {"ms_per_step": sum(s["solve_ms"] for s in stats) / len(stats),
                "cg_iterations": sum(s["iterations"] for s in stats) / len(stats),
                "initial_rel_residual": sum(s["initial_rel_residual"] for s in stats) / len(stats),
                "rel_residual": max(s["rel_residual"] for s in stats),
                "learned_used": sum(s["learned_used"] for s in stats)}

    solve = {args.solver: summary(st)}
    kt64 = kt if args.solver == "fp64" else {}
    others = []
    if not args.no_compare and args.solver != "fp64":
        others.append("fp64")
        if have_weights and not solvers.is_learned(args.solver):
            others.append("learned_ml")  # the learned warm start in front of the same mixed solver
    if args.ablation:
        others += [v for v in ("fp64_zero", "mixed_jacobi", "mixed_ml", "mixed_ml_zero", "learned_ml",
                               "learned_mixed")
                   if v != args.solver and v not in others and (have_weights or "learned" not in v)]
    del main_c
    for v in others:
        ref_c = make_computer(cfg, v, args.tol, args.weights)
        ref_c._bench_dt = cfg["dt"]
        if args.warmup:
            run(ref_c, args.warmup, f"{v} warmup")
        _, stv = run(ref_c, args.steps, f"{v} timed")
        if v == "fp64":
            kt64 = profiled_table(ref_c, args.steps, v)
            replay["cg_spmv"] = ref_c.replay_spmv("cg_spmv", REPLAYS)
            log(f"  cg_spmv: {replay['cg_spmv'] * 1e3:.1f} us per launch ({REPLAYS} back-to-back launches)")
        solve[v] = summary(stv)
        del ref_c

    def roof(table, name):
        """Achieved GB/s of a kernel: its algorithmic bytes per launch over the
        mean launch duration of REPLAYS back-to-back launches between two HIP
        events (per-launch events inside the steps add the dispatch of every
        launch; that figure is kept as ms_per_launch_in_steps)."""
        d = table.get(name)
        if not d:
            return None
        ms = replay.get(name, d["ms_per_launch"])
        gbs = d["bytes_per_launch"] / ms / 1e6
        return {"kernel": name, "bound": "hbm", "achieved": round(gbs, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                "traffic": TRAFFIC.get(name) if args.workload == "dambreak_1m" else None,
                "bytes_per_launch": d["bytes_per_launch"],
                "ms_per_launch": ms, "ms_per_launch_in_steps": d["ms_per_launch"],
                "timing": f"mean of {REPLAYS} back-to-back launches on the solver's operands"
                          if name in replay else "HIP events around each launch in the steps"}

    roofline = roof(kt, dom)
    roofline_fp64 = roof(kt64, "cg_spmv")

    # a plain copy kernel of known size: the practical HBM ceiling beside the 8 TB/s peak
    from openmps_amd import device_copy_probe
    copy_bytes = 1 << 30
    copy_ms = device_copy_probe(copy_bytes)
    copy = {"bytes_moved": 2 * copy_bytes, "ms": copy_ms, "GBps": 2 * copy_bytes / copy_ms / 1e6}
    log(f"  copy kernel: {copy['GBps']:.0f} GB/s")

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(cfg, args.tol, args.cpu_budget)
        except Exception as e:  # the oracle is optional for the measurement itself
            log(f"  cpu baseline failed: {e}")

    value = N * args.steps / elapsed
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("f64+f32" if solvers.is_mixed(args.solver) else "f64") +
                 ("+bf16" if solvers.is_learned(args.solver) else ""),
        "data": "synthetic",
        "config": {"workload": args.workload, "particles": N, "fluid_particles": nfluid,
                   "l0": cfg["env"]["l0"], "dt": cfg["dt"], "r_e_by_l0": cfg["env"]["r_e_by_l0"],
                   "r_lap_by_l0": cfg["env"]["r_lap_by_l0"], "solver": args.solver,
                   "stop_rule": f"|b-Ax|/|b| <= {args.tol:g}", "parallelism": "single GPU"},
        "fluid_particle_steps_per_s": nfluid * args.steps / elapsed,
        "solve": solve,
        "solve_speedup_vs_fp64": (solve["fp64"]["ms_per_step"] / solve_ms
                                  if "fp64" in solve and args.solver != "fp64" else None),
        "solve_speedup_learned_ml_vs_fp64": (solve["fp64"]["ms_per_step"] / solve["learned_ml"]["ms_per_step"]
                                             if "fp64" in solve and "learned_ml" in solve else None),
        "roofline": roofline,
        "roofline_fp64_spmv": roofline_fp64,
        "kernels": {k: {"ms": round(v["ms_per_launch"], 4), "GBps": round(v["GBps"], 1),
                        "launches": v["launches"], "bytes_per_launch": round(v["bytes_per_launch"])}
                    for k, v in kt.items()},
        "kernels_fp64": {k: {"ms": round(v["ms_per_launch"], 4), "GBps": round(v["GBps"], 1),
                             "launches": v["launches"], "bytes_per_launch": round(v["bytes_per_launch"])}
                         for k, v in kt64.items() if k.startswith("cg_")},
        "copy_kernel": copy,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
