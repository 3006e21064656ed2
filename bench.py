#!/usr/bin/env python3
"""Benchmark of the MPS time step (Computer::ForwardTime) on one MI355X.

Workload (default): the 1M-particle dam break, a 2.0 x 1.2 m tank with a
0.5 x 1.0 m water column at l0 = 0.7 mm (1,020,306 fluid + walls), positions
jittered +/-0.02 l0 from a fixed seed, r_e = 2.1 l0, Laplacian radius 3.1 l0,
dt = 1e-4 s, pressure solved to |b - A p| <= 1e-8 |b| (FP64 residual).

A step is one full ForwardTime: neighbour search, two density passes, HS/ECS,
viscosity + gravity, PPE assembly, the pressure solve, pressure gradient and
DS.  The headline solver (mixed_ml) steps the scene from rest: steps W+1..W+K
are the early window; it then runs untimed to step --advance (500: t = 0.05 s,
the collapsing column), and W warm-up + K timed steps from that developed
state give `value`.  The plain FP64 CG (the reference algorithm), the learned
warm start and the zero guess each start from the same developed state in
contexts of their own (W + K steps), as does the headline solver on the state
with the particles in a random order.  The CPU oracle runs --cpu-steps steps
from that state too: its time is the CPU baseline, and its states give the
error of the GPU steps against an FP64 CPU run.  Prints ONE JSON line on
stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s @1M; solve ms/step learned+mixed vs plain FP64 @res 1e-8"
HBM_PEAK_GBS = 8000.0
REPLAYS = 50
# untimed steps from rest before the timed window: the flow at t = 0.05 s
ADVANCE = 500
# HBM bytes per launch of the dambreak_1m SpMVs in the timed window (the
# developed flow) measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (one
# run each; FETCH_SIZE scaled by 2, WRITE_SIZE by 1: what the counters read for
# a copy kernel of known size), recorded with their commands in
# profiles/dambreak_1m_developed_kernels.md
TRAFFIC = {"cg_spmv32": 122.0e6, "cg_spmv": 372.3e6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_computer(cfg, solver: str, tol: float, weights: str | None = None):
    from openmps_amd import Computer, make_environment, solvers
    c = Computer(cfg["eps"], make_environment(cfg["env"]))
    solvers.configure(c, solver, tol, weights)
    c.add_particles(cfg["particles"])
    return c


def run(c, steps: int, label: str, every: int = 1):
    """Runs `steps` ForwardTime calls; returns wall seconds and per-step stats.
    Every call is synchronous at return (stream synchronised inside)."""
    stats = []
    t0 = time.perf_counter()
    for k in range(steps):
        c.forward_time(c._bench_dt)
        s = c.solve_stats()
        stats.append(s)
        if (k + 1) % every == 0 or k + 1 == steps:
            log(f"  [{label}] step {k + 1}/{steps}: CG {s['iterations']} it "
                f"({s['outer_iterations']} sweeps, precond {s.get('precond_used', 0)}), solve {s['solve_ms']:.1f} ms, "
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


def cpu_baseline(cfg, tol: float, state: dict, t: float, dt: float, steps: int, step0: int,
                 gpu64: dict, gpu_mixed_1: dict | None, solver: str):
    """The CPU oracle (C++/OpenMP FP64 restatement of the reference path, FP64
    CG with the same stop rule) for `steps` ForwardTime calls from the same
    developed state as the GPU comparisons, timed; its first and last states
    compared with the GPU's plain FP64 run (the reference algorithm) and its
    first with the headline solver's first step from that state."""
    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = oracle.Oracle(cfg["env"], cfg["eps"])
    o.set_stop_rule(1, tol)
    o.add_particles(state)
    o.set_time(t, dt)
    n = o.size()
    l0 = cfg["env"]["l0"]
    times, iters, errors = [], [], {}
    for k in range(1, steps + 1):
        t0 = time.perf_counter()
        o.forward_time(cfg["dt"])
        times.append(time.perf_counter() - t0)
        iters.append(o.cg_stats()["iterations"])
        log(f"  [cpu oracle] step {k}/{steps} from the state: {times[-1]:.2f} s, {iters[-1]} CG iterations")
        if k in gpu64:
            cnt, ids = o.neighbors()
            oc = {"counts": cnt, "ids": ids, "P": o.particles(), "b": o.aux("ppe_b")}
            errors[f"fp64_step{k}"] = compare_to_cpu(gpu64[k], oc, l0)
            if k == 1 and gpu_mixed_1 is not None:
                errors[f"{solver}_step1"] = compare_to_cpu(gpu_mixed_1, oc, l0)
    el = sum(times)
    return {"value": n * steps / el, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} ForwardTime steps of the same {n}-particle scene from the GPU run's state "
                      f"after {step0} steps (t = {t:.4f} s), FP64 CG to the same stop rule "
                      f"({sum(iters) / len(iters):.0f} iterations a step), {el:.1f} s",
            "ms_per_step": el / steps * 1e3, "cg_iterations": iters,
            "cpu_model": cpu_model(), "nproc": os.cpu_count(), "omp_threads": threads,
            "error": errors}


def run_batch(args) -> int:
    """64 uneven independent scenes (scenes.batch: dam breaks and static
    columns, 19k-207k particles, l0 0.5-2 mm): the batch API steps them
    together (per-scene stages from `threads` host threads, the solves of all
    scenes as grouped launches, one MLP inference over all rows), each scene on
    its own CFL step.  value = total particle-steps per second over the K timed
    steps.  Beside it, from the same initial state: the same scenes as 64
    separate contexts stepped from the same host threads (each scene's solve
    launched on its own), and every scene alone with the plain FP64 CG for the
    W + K steps, whose final state gives the batch's per-scene error."""
    from concurrent.futures import ThreadPoolExecutor
    from openmps_amd import Batch, learned, scenes, solvers
    cfgs = scenes.batch()
    solver = args.solver or "learned_amg"
    model = learned.Model.load(args.weights or learned.DEFAULT_WEIGHTS) if solvers.is_learned(solver) else None
    log(f"building {len(cfgs)} scenes ...")
    sizes = [len(cfg["particles"]["x"]) for cfg in cfgs]
    total = sum(sizes)
    W, K = args.warmup, args.steps

    def contexts(preset):
        out = []
        for cfg in cfgs:
            c = make_computer(cfg, preset, args.tol, args.weights) if model is None or not solvers.is_learned(preset) \
                else _with_model(cfg, preset, args.tol, model)
            out.append(c)
        return out

    # 1. the batch
    comps = contexts(solver)
    batch = Batch(comps, host_threads=args.threads)
    for k in range(W):
        batch.forward_time()
    log(f"  batch: {len(comps)} scenes, {total} particles, {args.threads} host threads: timing {K} steps")
    solve_ms, launches = [], []
    t0 = time.perf_counter()
    for k in range(K):
        batch.forward_time()
        info = batch.solve_info()
        solve_ms.append(info["solve_ms"])
        launches.append(info["launches"])
    elapsed = time.perf_counter() - t0
    stats = [c.solve_stats() for c in comps]
    final = [c.particles() for c in comps]
    log(f"  batch: {elapsed / K * 1e3:.1f} ms/step, solve {np.mean(solve_ms):.1f} ms, "
        f"{np.mean(launches):.0f} grouped launches a solve")
    del batch

    # 2. the same scenes as separate contexts, each stepping (and solving) on its own
    sep = contexts(solver)
    groups = [[] for _ in range(args.threads)]
    load = [0] * args.threads
    for k in sorted(range(len(sep)), key=lambda k: -sizes[k]):  # largest first onto the lightest share
        w = load.index(min(load))
        groups[w].append(k)
        load[w] += sizes[k]

    def work(idx, steps):
        for _ in range(steps):
            for k in idx:
                sep[k].forward_time()
        return len(idx)

    with ThreadPoolExecutor(max_workers=args.threads) as ex:
        list(ex.map(lambda g: work(g, W), groups))
        t1 = time.perf_counter()
        list(ex.map(lambda g: work(g, K), groups))
        el_sep = time.perf_counter() - t1
    log(f"  separate contexts: {el_sep / K * 1e3:.1f} ms/step")
    del sep

    # 3. every scene alone with the FP64 CG (the reference algorithm), W + K
    # steps; and, as a control of what the stop rule itself resolves, alone
    # with the FP64 CG to a 100 times tighter residual
    def err(P, Q, l0):
        fl = Q["type"] == 0
        return (float(np.linalg.norm(P["p"][fl] - Q["p"][fl]) / np.linalg.norm(Q["p"][fl])),
                float(max(np.abs(P["x"] - Q["x"]).max(), np.abs(P["z"] - Q["z"]).max()) / l0))

    errs = []
    t2 = time.perf_counter()
    for s, cfg in enumerate(cfgs):
        runs = []
        for tol in (args.tol, args.tol / 100):
            o = make_computer(cfg, "fp64", tol)
            for _ in range(W + K):
                o.forward_time()
            runs.append(o.particles())
            del o
        pe_, xe_ = err(final[s], runs[0], cfg["l0"])
        pc_, xc_ = err(runs[1], runs[0], cfg["l0"])
        errs.append({"scene": s, "kind": cfg["kind"], "particles": sizes[s], "l0": cfg["l0"],
                     "p_rel_l2": pe_, "position_max_diff_in_l0": xe_,
                     "control_fp64_tight_p_rel_l2": pc_, "control_fp64_tight_position_max_diff_in_l0": xc_,
                     "cg_iterations": stats[s]["iterations"],
                     "learned_used": stats[s]["learned_used"], "rel_residual": stats[s]["rel_residual"]})
    log(f"  FP64 runs alone: {time.perf_counter() - t2:.1f} s")
    pe = [e["p_rel_l2"] for e in errs]
    xe = [e["position_max_diff_in_l0"] for e in errs]
    pc = [e["control_fp64_tight_p_rel_l2"] for e in errs]
    out = {"metric": "particle-steps/s, 64 uneven scenes batched on one GPU", "value": total * K / elapsed,
           "unit": "particle-steps/s", "n_gpus": 1, "steps": K, "warmup": W,
           "ms_per_step": elapsed / K * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": ("f64+f32" if solvers.is_mixed(solver) else "f64") +
                                         ("+bf16" if solvers.is_learned(solver) else ""),
           "data": "synthetic",
           "config": {"workload": "batch64", "scenes": len(cfgs), "particles": total,
                      "particles_min_max": [min(sizes), max(sizes)],
                      "kinds": {k: sum(1 for c in cfgs if c["kind"] == k) for k in ("dambreak", "static")},
                      "solver": solver, "host_threads": args.threads,
                      "stop_rule": f"|b-Ax|/|b| <= {args.tol:g}", "dt": "each scene's CFL step, at most 1e-4 s",
                      "timed_steps": f"{W + 1}..{W + K} from rest"},
           "solve_ms_per_step": float(np.mean(solve_ms)), "grouped_launches_per_solve": float(np.mean(launches)),
           "separate_contexts": {"particle_steps_per_s": total * K / el_sep, "ms_per_step": el_sep / K * 1e3,
                                 "host_threads": args.threads},
           "speedup_vs_separate_contexts": el_sep / elapsed,
           "error_vs_fp64_alone": {
               "after_steps": W + K, "p_rel_l2_max": max(pe), "p_rel_l2_median": float(np.median(pe)),
               "position_max_diff_in_l0_max": max(xe),
               "by_kind": {k: {"p_rel_l2_max": max(e["p_rel_l2"] for e in errs if e["kind"] == k),
                               "control_fp64_tight_p_rel_l2_max": max(e["control_fp64_tight_p_rel_l2"]
                                                                      for e in errs if e["kind"] == k)}
                           for k in ("dambreak", "static")},
               "control_fp64_tight_p_rel_l2_median": float(np.median(pc)),
               "per_scene": errs},
           "max_rel_residual": max(s_["rel_residual"] for s_ in stats),
           "learned_used_last_step": sum(s_["learned_used"] for s_ in stats)}
    print(json.dumps(out), flush=True)
    return 0


def _with_model(cfg, preset, tol, model):
    from openmps_amd import Computer, make_environment, solvers
    c = Computer(cfg["eps"], make_environment(cfg["env"]))
    solvers.configure(c, preset, tol, model=model)
    c.add_particles(cfg["particles"])
    return c


def summary(stats, elapsed=None):
    out = {"solve_ms_per_step": sum(s["solve_ms"] for s in stats) / len(stats),
           "cg_iterations": sum(s["iterations"] for s in stats) / len(stats),
           "initial_rel_residual": sum(s["initial_rel_residual"] for s in stats) / len(stats),
           "rel_residual": max(s["rel_residual"] for s in stats),
           "learned_used": sum(s["learned_used"] for s in stats)}
    names = {0: "none", 1: "jacobi", 2: "multilevel", 3: "amg"}
    used = [names.get(s.get("precond_used", 0), "?") for s in stats]
    if any(u in ("multilevel", "amg") for u in used):
        out["preconditioner_steps"] = {u: used.count(u) for u in sorted(set(used))}
    if elapsed is not None:
        out["step_ms"] = elapsed / len(stats) * 1e3
    return out


def from_state(cfg, solver: str, tol: float, state: dict, t: float, dt: float, weights=None):
    """A context holding `state` (particles as Particles() returned them) at
    time t: the other solvers start from the headline run's developed state."""
    from openmps_amd import Computer, make_environment, solvers
    c = Computer(cfg["eps"], make_environment(cfg["env"]))
    solvers.configure(c, solver, tol, weights)
    c.add_particles(state)
    c.set_time(t, dt)
    c._bench_dt = cfg["dt"]
    return c


def capture(c):
    """Lists, particles and the assembled right-hand side after a step."""
    cnt, ids = c.neighbors()
    return {"counts": cnt, "ids": ids, "P": c.particles(), "b": c.aux("ppe_b")}


def compare_to_cpu(g, o, l0):
    """Error of a GPU step against the oracle's step from the same state."""
    P, Q = g["P"], o["P"]
    fl = Q["type"] == 0
    same_lists = bool(np.array_equal(g["counts"], o["counts"]) and np.array_equal(g["ids"], o["ids"]))
    dn = np.abs(P["n"] - Q["n"])
    return {"neighbor_lists_equal": same_lists,
            "n_bit_equal": bool(np.array_equal(P["n"], Q["n"])),
            "n_max_rel_diff": float((dn / np.maximum(np.abs(Q["n"]), 1e-300)).max()),
            "ppe_b_bit_equal": bool(np.array_equal(g["b"], o["b"])),
            "p_rel_l2": float(np.linalg.norm(P["p"][fl] - Q["p"][fl]) / np.linalg.norm(Q["p"][fl])),
            "position_max_diff_in_l0": float(max(np.abs(P["x"] - Q["x"]).max(), np.abs(P["z"] - Q["z"]).max()) / l0),
            "velocity_rel_l2": float(np.sqrt(np.sum((P["u"] - Q["u"]) ** 2 + (P["w"] - Q["w"]) ** 2) /
                                             np.sum(Q["u"] ** 2 + Q["w"] ** 2)))}


def cpu_model() -> str:
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--advance", type=int, default=ADVANCE,
                    help="untimed steps from rest before the timed window (the developed flow)")
    ap.add_argument("--workload", default="dambreak_1m")
    from openmps_amd import solvers
    ap.add_argument("--solver", default=None, choices=sorted(solvers.PRESETS),
                    help="headline solver (default: mixed_auto, FP32 CG with the faster of the "
                         "multilevel and the cell-grid multigrid preconditioners inside FP64 refinement; "
                         "learned_amg, the multigrid one with the learned warm start, for batch64)")
    ap.add_argument("--weights", default=None, help="warm-start MLP weights (.npz)")
    ap.add_argument("--ablation", action="store_true", help="also time further solver presets")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-compare", action="store_true", help="skip the other solvers")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=3, help="oracle steps from the developed state")
    ap.add_argument("--threads", type=int, default=8, help="host threads of the batch64 workload")
    ap.add_argument("--save-state", default=None, help="write the developed state (.npz) after the advance")
    ap.add_argument("--from-state", default=None,
                    help="start from a state written by --save-state (no early window, no advance)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        return 2
    if args.workload == "batch64":
        return run_batch(args)
    args.solver = args.solver or "mixed_auto"
    from openmps_amd import scenes
    from openmps_amd import learned as _learned
    have_weights = os.path.exists(args.weights or _learned.DEFAULT_WEIGHTS)
    log(f"building scene {args.workload} ...")
    cfg = scenes.config(args.workload)
    N = len(cfg["particles"]["x"])
    nfluid = int((cfg["particles"]["type"] == 0).sum())
    l0 = cfg["env"]["l0"]
    log(f"  {N} particles ({nfluid} fluid)")
    W, K = args.warmup, args.steps

    early = {}
    if args.from_state:
        # a developed state saved by an earlier run (--save-state)
        z = np.load(args.from_state)
        state = {k: z[k] for k in ("x", "z", "u", "w", "p", "n", "type")}
        t_state, dt_state, step0 = float(z["t"]), float(z["dt"]), int(z["step"])
        main_c = from_state(cfg, args.solver, args.tol, state, t_state, dt_state, args.weights)
        log(f"  state after {step0} steps from {args.from_state}, t = {t_state:.4f} s")
    else:
        # 1. the early window: the headline solver's steps W+1 .. W+K from rest
        main_c = make_computer(cfg, args.solver, args.tol, args.weights)
        main_c._bench_dt = cfg["dt"]
        run(main_c, W, f"{args.solver} early warmup")
        el, st = run(main_c, K, f"{args.solver} early timed")
        early[args.solver] = summary(st, el)
        # 2. the developed flow: untimed steps up to `advance`, then the state the
        # other solvers start from
        rest = max(0, args.advance - W - K)
        run(main_c, rest, f"{args.solver} advance", every=50)
        state = main_c.particles()
        t_state, dt_state = main_c.time()
        step0 = W + K + rest
        log(f"  developed state after {step0} steps, t = {t_state:.4f} s")
        if args.save_state:
            np.savez(args.save_state, t=t_state, dt=dt_state, step=step0, **state)
    # 3. headline window from the developed state: W warm-up + K timed steps
    el_w, st_w = run(main_c, 1, f"{args.solver} warmup")
    gpu_mixed_1 = capture(main_c) if not args.no_cpu_baseline else None
    run(main_c, W - 1, f"{args.solver} warmup")
    elapsed, st = run(main_c, K, f"{args.solver} timed")
    kt = profiled_table(main_c, K, args.solver)
    dom = {"none": "cg_iter32", "jacobi": "cg_iter32", "multilevel": "cg_spmv32", "amg": "cg_spmv32",
           "auto": "cg_spmv32"}[
        solvers.preconditioner(args.solver)] if solvers.is_mixed(args.solver) else "cg_spmv"
    replay = {}
    if dom in ("cg_spmv32", "cg_spmv"):
        replay[dom] = main_c.replay_spmv(dom, REPLAYS)
        log(f"  {dom}: {replay[dom] * 1e3:.1f} us per launch ({REPLAYS} back-to-back launches)")
    solve = {args.solver: summary(st, elapsed)}
    del main_c

    # 4. the other solvers from the same developed state
    kt64 = kt if args.solver == "fp64" else {}
    gpu64 = {}
    others = []
    prec = solvers.preconditioner(args.solver)
    learned_v = {"amg": "learned_amg", "multilevel": "learned_ml", "jacobi": "learned_mixed",
                 "auto": "learned_auto"}.get(prec)
    zero_v = {"amg": "mixed_amg_zero", "multilevel": "mixed_ml_zero", "auto": "mixed_auto_zero"}.get(prec)
    if not args.no_compare:
        if args.solver != "fp64":
            others.append("fp64")
        if have_weights and learned_v and not solvers.is_learned(args.solver):
            others.append(learned_v)  # the learned warm start ...
        if zero_v and args.solver != zero_v:
            others.append(zero_v)  # ... against a zero guess (and the previous-step one: the headline)
        for v in ("mixed_ml", "mixed_amg"):  # the other preconditioner
            if v != args.solver and v not in others:
                others.append(v)
    if args.ablation:
        others += [v for v in ("fp64_zero", "mixed_jacobi", "mixed_ml", "learned_mixed")
                   if v != args.solver and v not in others and (have_weights or "learned" not in v)]
    for v in others:
        c = from_state(cfg, v, args.tol, state, t_state, dt_state, args.weights)
        if v == "fp64" and not args.no_cpu_baseline:
            # the reference algorithm's first steps from the state, kept for the CPU comparison
            for k in range(max(1, args.cpu_steps)):
                run(c, 1, f"{v} from state")
                if k == 0 or k == args.cpu_steps - 1:
                    gpu64[k + 1] = capture(c)
            run(c, max(0, W - max(1, args.cpu_steps)), f"{v} warmup")
        else:
            run(c, W, f"{v} warmup")
        el_v, stv = run(c, K, f"{v} timed")
        if v == "fp64":
            kt64 = profiled_table(c, K, v)
            replay["cg_spmv"] = c.replay_spmv("cg_spmv", REPLAYS)
            log(f"  cg_spmv: {replay['cg_spmv'] * 1e3:.1f} us per launch ({REPLAYS} back-to-back launches)")
        solve[v] = summary(stv, el_v)
        del c
    # the headline solver on the same state with the particles in a random order:
    # the neighbour walks and the FP32 operator's 16-bit column offsets depend on index locality
    shuffled = None
    if not args.no_compare:
        perm = np.random.Generator(np.random.PCG64(3)).permutation(N)
        c = from_state(cfg, args.solver, args.tol, {k: v[perm] for k, v in state.items()}, t_state, dt_state,
                       args.weights)
        run(c, W, f"{args.solver} shuffled warmup")
        el_s, sts = run(c, K, f"{args.solver} shuffled timed")
        shuffled = summary(sts, el_s)
        shuffled["particle_steps_per_s"] = N * K / el_s
        del c
    # the early window of the reference algorithm, for the speedup at rest
    if not args.no_compare and args.solver != "fp64" and not args.from_state:
        c = make_computer(cfg, "fp64", args.tol)
        c._bench_dt = cfg["dt"]
        run(c, W, "fp64 early warmup")
        el_e, ste = run(c, K, "fp64 early timed")
        early["fp64"] = summary(ste, el_e)
        del c

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
            cpu = cpu_baseline(cfg, args.tol, state, t_state, dt_state, max(1, args.cpu_steps), step0,
                               gpu64, gpu_mixed_1, args.solver)
        except Exception as e:  # the oracle is optional for the measurement itself
            log(f"  cpu baseline failed: {e}")

    value = N * K / elapsed
    sp = lambda a, b: (solve[a]["solve_ms_per_step"] / solve[b]["solve_ms_per_step"]
                       if a in solve and b in solve else None)
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("f64+f32" if solvers.is_mixed(args.solver) else "f64") +
                 ("+bf16" if solvers.is_learned(args.solver) else ""),
        "data": "synthetic",
        "config": {"workload": args.workload, "particles": N, "fluid_particles": nfluid,
                   "l0": l0, "dt": cfg["dt"], "r_e_by_l0": cfg["env"]["r_e_by_l0"],
                   "r_lap_by_l0": cfg["env"]["r_lap_by_l0"], "solver": args.solver,
                   "timed_steps": f"{step0 + W + 1}..{step0 + W + K} (t = {t_state:.4f} s at step {step0})",
                   "stop_rule": f"|b-Ax|/|b| <= {args.tol:g}", "parallelism": "single GPU"},
        "fluid_particle_steps_per_s": nfluid * K / elapsed,
        "fp64_particle_steps_per_s": (N / (solve["fp64"]["step_ms"] * 1e-3) if "fp64" in solve else None),
        "solve": solve,
        "solve_speedup_vs_fp64": sp("fp64", args.solver),
        "learned_solver": learned_v,
        "solve_speedup_learned_vs_fp64": sp("fp64", learned_v) if learned_v else None,
        # > 1: the learned start is faster than the previous-step guess of the same solver
        "learned_vs_previous_guess_solve_ratio": (sp(args.solver, learned_v)
                                                  if learned_v and not solvers.is_learned(args.solver) else None),
        "solve_speedup_vs_mixed_ml": sp("mixed_ml", args.solver) if args.solver != "mixed_ml" else None,
        "shuffled_order": shuffled,
        "early_window": {"steps": f"{W + 1}..{W + K} from rest", "solve": early,
                         "solve_speedup_vs_fp64": (early["fp64"]["solve_ms_per_step"] /
                                                   early[args.solver]["solve_ms_per_step"]
                                                   if "fp64" in early and args.solver in early else None),
                         "particle_steps_per_s": N / (early[args.solver]["step_ms"] * 1e-3)}
                        if args.solver in early else None,
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
