"""Validation runs on the GPU path.

hydrostatic  A still water column (the static-pressure case): the fluid
             pressure against rho g (H - z), H the free-surface height (half a
             spacing above the top layer), for the last step and for the
             pressure averaged over the final part of the run.  Reported: the
             relative L2 error over the fluid particles deeper than `skin`
             spacings below the surface, and the largest error of the
             pressure averaged over horizontal bands 10 spacings thick
             (relative to the bottom pressure rho g H).

dambreak     The collapsing water column (column width L, height 2L, tank
             4L: the Koshizuka & Oka 1996 geometry): the surge front (largest
             x of the fluid particles, as Benchmark/DamBreak/
             koshizukaoka1996_edge.py:15-20 takes it) over t sqrt(2g/L),
             against the experiments and computations shipped in
             openmps_amd/data/dambreak (front / L at their sample times).

compare      The same scene stepped with two solver presets (e.g. mixed_ml
             against fp64) for a number of steps; the relative L2 difference of
             the pressure and the RMS position difference, in spacings.

Usage:
  python -m openmps_amd.validate hydrostatic --scene static_pressure --end-time 0.5 --solver fp64
  python -m openmps_amd.validate hydrostatic --scene still_1m --steps 500 --solver mixed_ml
  python -m openmps_amd.validate compare --scene dambreak_1m --steps 500 --solvers mixed_ml fp64
  python -m openmps_amd.validate dambreak --scene dambreak_125k --t-star 3.0 --solver mixed_ml
Each run prints one JSON line on stdout (progress on stderr).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def _computer(cfg, preset: str, tol: float):
    from . import Computer, make_environment, solvers
    c = Computer(cfg["eps"], make_environment(cfg["env"]))
    solvers.configure(c, preset, tol)
    c.add_particles(cfg["particles"])
    return c


def _step(c, cfg, steps: int, label: str, every: int = 20, end_time: float | None = None, on_step=None):
    """ForwardTime() calls with dt from the CFL rule (Computer.hpp:759-777)
    capped by the scene's maximum step: `steps` of them, or until the time
    reaches `end_time`.  `on_step(c)` runs after each.  Returns (steps taken,
    mean CG iterations)."""
    t0 = time.perf_counter()
    iters = 0
    k = 0
    while (end_time is None and k < steps) or (end_time is not None and c.time()[0] < end_time - 1e-12):
        c.forward_time()
        k += 1
        s = c.solve_stats()
        iters += s["iterations"]
        if on_step is not None:
            on_step(c)
        if k % every == 0:
            _log(f"  [{label}] step {k}, t = {c.time()[0]:.5f} s, {time.perf_counter() - t0:.1f} s, "
                 f"{s['iterations']} CG iterations, |b-Ax|/|b| {s['rel_residual']:.1e}")
    return k, iters / max(k, 1)


def hydrostatic_error(P: dict, cfg: dict, skin: float = 5.0, band: float = 10.0) -> dict:
    """Pressure of the fluid particles against rho g (H - z), H the free
    surface (half a spacing above the initial top layer)."""
    env = cfg["env"]
    l0, rho, g = env["l0"], env["rho"], env["g"]
    fluid = P["type"] == 0
    z0 = cfg["particles"]["z"][cfg["particles"]["type"] == 0]
    H = z0.max() + 0.5 * l0
    z, p = P["z"][fluid], P["p"][fluid]
    deep = z < H - skin * l0
    want = rho * g * (H - z)
    rel_l2 = float(np.linalg.norm(p[deep] - want[deep]) / np.linalg.norm(want[deep]))
    edges = np.arange(0.0, H + band * l0, band * l0)
    idx = np.digitize(z, edges)
    worst = 0.0
    for b in np.unique(idx[deep]):
        m = (idx == b) & deep
        if m.sum() < 10:
            continue
        worst = max(worst, abs(p[m].mean() - want[m].mean()) / (rho * g * H))
    return {"rel_l2_deep": rel_l2, "band_mean_max_err_rel_bottom": worst, "particles": int(deep.sum()),
            "surface_height_m": float(H)}


def run_hydrostatic(scene: str, steps: int, preset: str, tol: float, end_time: float | None = None,
                    average_from: float = 0.5) -> dict:
    """Steps the scene (to `end_time` if given, else `steps` steps) and
    compares the last pressure and the pressure averaged over the run's
    final part (from average_from * the end, per particle, with its mean
    height) with rho g (H - z).  MPS pressure oscillates from step to step
    about the hydrostatic value, so the average is the meaningful check."""
    from . import scenes
    cfg = scenes.config(scene)
    c = _computer(cfg, preset, tol)
    acc = {"p": None, "z": None, "k": 0}
    t_avg = average_from * end_time if end_time is not None else None
    k_avg = int(average_from * steps)
    count = [0]

    def on_step(cc):
        count[0] += 1
        if (t_avg is not None and cc.time()[0] >= t_avg) or (t_avg is None and count[0] > k_avg):
            P = cc.particles()
            if acc["p"] is None:
                acc["p"], acc["z"] = np.zeros_like(P["p"]), np.zeros_like(P["z"])
            acc["p"] += P["p"]
            acc["z"] += P["z"]
            acc["k"] += 1

    k, its = _step(c, cfg, steps, preset, every=max(20, steps // 25) if end_time is None else 1000,
                   end_time=end_time, on_step=on_step)
    P = c.particles()
    out = {"check": "hydrostatic", "scene": scene, "steps": k, "solver": preset, "time_s": c.time()[0],
           "mean_cg_iterations": its, "last": hydrostatic_error(P, cfg)}
    if acc["k"]:
        Q = dict(P)
        Q["p"], Q["z"] = acc["p"] / acc["k"], acc["z"] / acc["k"]
        out["averaged"] = hydrostatic_error(Q, cfg)
        out["averaged_steps"] = acc["k"]
    return out


DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "dambreak")


def front_references() -> dict:
    """name -> (t*, Z/L) arrays of the shipped surge-front data."""
    out = {}
    for f in sorted(os.listdir(DATA)):
        if f.endswith(".csv"):
            a = np.loadtxt(os.path.join(DATA, f), delimiter=",", skiprows=1, ndmin=2)
            out[f[:-4]] = (a[:, 0], a[:, 1])
    return out


def run_dambreak(scene: str, t_star: float, preset: str, tol: float, every: float = 0.005) -> dict:
    """Front position Z/L against t sqrt(2g/L) up to t_star, sampled every
    `every` seconds, compared with every reference curve at its own sample
    times inside the simulated range."""
    from . import scenes
    cfg = scenes.config(scene)
    P0 = cfg["particles"]
    fluid0 = P0["type"] == 0
    l0, g = cfg["env"]["l0"], cfg["env"]["g"]
    L = P0["x"][fluid0].max() - P0["x"][fluid0].min() + l0  # column width
    scale = np.sqrt(2 * g / L)
    c = _computer(cfg, preset, tol)
    ts, zs = [0.0], [P0["x"][fluid0].max() / L]
    end = t_star / scale
    nxt = [every]

    def on_step(cc):
        t = cc.time()[0]
        if t >= nxt[0] - 1e-12:
            nxt[0] += every
            P = cc.particles()
            ts.append(t * scale)
            zs.append(P["x"][P["type"] == 0].max() / L)

    k, its = _step(c, cfg, 0, preset, every=500, end_time=end, on_step=on_step)
    ts, zs = np.array(ts), np.array(zs)
    cmp = {}
    for name, (rt, rz) in front_references().items():
        m = (rt >= ts[0]) & (rt <= ts[-1])
        if m.sum() == 0:
            continue
        ours = np.interp(rt[m], ts, zs)
        d = ours - rz[m]
        cmp[name] = {"points": int(m.sum()), "mean_diff": float(d.mean()), "max_abs_diff": float(np.abs(d).max()),
                     "rms_rel": float(np.sqrt(np.mean((d / rz[m]) ** 2)))}
    return {"check": "dambreak", "scene": scene, "solver": preset, "steps": k, "mean_cg_iterations": its,
            "column_width_m": float(L), "t_star_end": float(ts[-1]),
            "front": {"t_star": [round(float(v), 4) for v in ts], "z_over_L": [round(float(v), 4) for v in zs]},
            "versus": cmp}


def run_compare(scene: str, steps: int, presets: list[str], tol: float) -> dict:
    from . import scenes
    cfg = scenes.config(scene)
    states = []
    for pr in presets:
        c = _computer(cfg, pr, tol)
        _, its = _step(c, cfg, steps, pr)
        P = c.particles()
        states.append((pr, P, its, c.time()[0]))
        del c
    (a, Pa, ia, _), (b, Pb, ib, _) = states[0], states[1]
    fl = Pb["type"] == 0
    dp = float(np.linalg.norm(Pa["p"][fl] - Pb["p"][fl]) / np.linalg.norm(Pb["p"][fl]))
    dx = np.sqrt(((Pa["x"][fl] - Pb["x"][fl]) ** 2 + (Pa["z"][fl] - Pb["z"][fl]) ** 2).mean())
    return {"check": "compare", "scene": scene, "steps": steps, "solvers": presets,
            "time_s": states[1][3],
            "pressure_rel_l2": dp, "position_rms_in_l0": float(dx / cfg["env"]["l0"]),
            "mean_cg_iterations": {a: ia, b: ib}}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = ap.add_subparsers(dest="cmd", required=True)
    h = sub.add_parser("hydrostatic")
    h.add_argument("--scene", default="still_1m")
    h.add_argument("--steps", type=int, default=500)
    h.add_argument("--solver", default="mixed_ml")
    h.add_argument("--tol", type=float, default=1e-8)
    h.add_argument("--end-time", type=float, default=None, help="step to this time instead of --steps")
    h.add_argument("--average-from", type=float, default=0.5,
                   help="average the pressure from this fraction of the run on")
    c = sub.add_parser("compare")
    c.add_argument("--scene", default="dambreak_1m")
    c.add_argument("--steps", type=int, default=500)
    c.add_argument("--solvers", nargs=2, default=["mixed_ml", "fp64"])
    c.add_argument("--tol", type=float, default=1e-8)
    d = sub.add_parser("dambreak")
    d.add_argument("--scene", default="dambreak_125k")
    d.add_argument("--t-star", type=float, default=3.0)
    d.add_argument("--solver", default="mixed_ml")
    d.add_argument("--tol", type=float, default=1e-8)
    args = ap.parse_args(argv)
    if args.cmd == "dambreak":
        out = run_dambreak(args.scene, args.t_star, args.solver, args.tol)
    elif args.cmd == "hydrostatic":
        out = run_hydrostatic(args.scene, args.steps, args.solver, args.tol, args.end_time, args.average_from)
    else:
        out = run_compare(args.scene, args.steps, args.solvers, args.tol)
    if args.cmd == "dambreak":
        front = out["front"]
        _log("  t*   Z/L: " + ", ".join(f"{a:.2f} {b:.3f}" for a, b in
                                       list(zip(front["t_star"], front["z_over_L"]))[::10]))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
