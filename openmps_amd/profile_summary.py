"""Markdown summary of a `rocprofv3 --kernel-trace --stats` run of bench.py.

Joins the profiler's kernel dispatch durations with the algorithmic bytes per
launch that bench.py reports for the same command (its JSON line), giving the
achieved GB/s of every path kernel against the 8 TB/s HBM peak.  With the
databases of two `--pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs it adds the
measured HBM traffic of the solver kernels, scaled by what the same counters
read for the bench's calibration copy kernel (`k_copy_probe`, a known number
of bytes read and written).

  python -m openmps_amd.profile_summary --db trace.db --bench bench.json \
      --out profiles/x.md --title "..." --command "..." \
      [--fetch fetch.db --write write.db]
"""
from __future__ import annotations

import argparse
import json
import re
import sqlite3
from collections import defaultdict

import numpy as np

HBM_PEAK_GBS = 8000.0

# profiler kernel name (template arguments kept) -> kernel id of the library's
# own per-kernel statistics (mps.h MPS_K_*)
KERNEL_IDS = {
    "k_cell_keys": "cell_keys", "k_cell_scatter": "cell_scatter", "k_cell_sort": "cell_sort",
    "k_neighbors": "neighbors", "k_density<true, true, false>": "density_ecs",
    "k_density<true, false, true>": "density_save", "k_explicit": "explicit",
    "k_ppe_count": "ppe_count", "k_ppe_fill": "ppe_fill", "k_cg_init": "cg_init",
    "k_cg_spmv<false>": "cg_spmv", "k_cg_spmv<true>": "cg_spmv", "k_cg_update": "cg_update",
    "k_gradient": "gradient", "k_ds": "ds",
    "k_ml_spmv<false, mps::(anonymous namespace)::MatFree>": "cg_spmv32",
    "k_ml_spmv<true, mps::(anonymous namespace)::MatFree>": "cg_spmv32",
    "k_ml_spmv<false, mps::(anonymous namespace)::MatFreeRec>": "cg_spmv32",
    "k_ml_spmv<true, mps::(anonymous namespace)::MatFreeRec>": "cg_spmv32",
    "k_density_explicit": "density_ecs", "k_density_ppe_count": "density_save",
    "k_ml_prec": "ml_prec",
    "k_cgf_iter<mps::(anonymous namespace)::MatFree>": "cg_iter32",
    "k_features<unsigned char>": "features", "k_features<unsigned short>": "features",
}


def short_name(n: str) -> str:
    n = n.replace("void ", "").replace("mps::(anonymous namespace)::", "", 1)
    depth, out = 0, []
    for ch in n:  # cut the argument list, keep template arguments
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            break
        out.append(ch)
    return "".join(out).strip()


def kernel_durations(db: str) -> dict:
    con = sqlite3.connect(db)
    d = defaultdict(list)
    for name, du in con.execute("select name, duration from kernels"):
        d[short_name(name)].append(du / 1000.0)
    return {k: np.array(v) for k, v in d.items()}


def pmc_values(db: str, counter: str) -> dict:
    """Per-dispatch counter values by kernel name (rocprofv3 `pmc_events` view)."""
    con = sqlite3.connect(db)
    out = defaultdict(list)
    for name, v in con.execute("select name, counter_value from pmc_events where counter_name = ?",
                               (counter,)):
        out[short_name(name)].append(float(v))
    return {k: np.array(v) for k, v in out.items()}


def working(a: np.ndarray) -> np.ndarray:
    """Launches longer than 10 % of the kernel's longest: solver launches after
    convergence exit at once and are left out."""
    return a[a > 0.1 * a.max()] if len(a) else a


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--db", required=True)
    ap.add_argument("--bench", required=True, help="bench.py JSON line of the profiled command")
    ap.add_argument("--out", required=True)
    ap.add_argument("--title", required=True)
    ap.add_argument("--command", required=True)
    ap.add_argument("--note", default="")
    ap.add_argument("--fetch", help="database of a --pmc FETCH_SIZE run")
    ap.add_argument("--write", help="database of a --pmc WRITE_SIZE run")
    args = ap.parse_args(argv)

    dur = kernel_durations(args.db)
    bench = json.loads(open(args.bench).read().strip().splitlines()[-1])
    table = {}
    for src in ("kernels", "kernels_fp64"):
        for k, v in bench.get(src, {}).items():
            table.setdefault(k, v)
    total = sum(a.sum() for a in dur.values())
    lines = [f"# {args.title}", "", "Command (from the repository root, one MI355X):", "", "```",
             args.command, "```", ""]
    if args.note:
        lines += [args.note, ""]
    lines += ["Durations: rocprofv3 kernel dispatch records.  'working' = launches longer than 10 % of",
              "the kernel's longest (solver launches after convergence exit at once).  Bytes: the",
              "algorithmic bytes per launch the library accounts for the kernel (each array once per",
              "launch), from the bench JSON of the same command; GB/s = bytes / mean working duration.",
              "", "| kernel | calls | total ms | % time | mean working us | MB / launch | GB/s | % of 8 TB/s |",
              "|---|---|---|---|---|---|---|---|"]
    for k in sorted(dur, key=lambda k: -dur[k].sum()):
        a = dur[k]
        w = working(a)
        kid = KERNEL_IDS.get(k)
        mb = gbs = frac = "—"
        if kid and kid in table and "bytes_per_launch" in table[kid]:
            bpl = table[kid]["bytes_per_launch"]
            mb = f"{bpl / 1e6:.1f}"
            g = bpl / (w.mean() * 1e3)
            gbs, frac = f"{g:.0f}", f"{100 * g / HBM_PEAK_GBS:.1f}"
        lines.append(f"| `{k}` | {len(a)} | {a.sum() / 1e3:.3f} | {100 * a.sum() / total:.1f} | "
                     f"{w.mean():.2f} | {mb} | {gbs} | {frac} |")
    if args.fetch and args.write:
        f = pmc_values(args.fetch, "FETCH_SIZE")
        wr = pmc_values(args.write, "WRITE_SIZE")
        lines += ["", "## Measured HBM traffic (PMC)", ""]
        cal = None
        if "k_copy_probe" in f and "k_copy_probe" in wr:
            nbytes = bench.get("copy_kernel", {}).get("bytes_moved", 0) / 2
            cal = (nbytes / (np.median(f["k_copy_probe"]) * 1024), nbytes / (np.median(wr["k_copy_probe"]) * 1024))
            lines.append(f"Calibration: the copy kernel reads and writes {nbytes / 2**20:.0f} MiB; "
                         f"FETCH_SIZE reads {np.median(f['k_copy_probe']) / 1024:.0f} MiB, WRITE_SIZE "
                         f"{np.median(wr['k_copy_probe']) / 1024:.0f} MiB (scale factors {cal[0]:.3f}, {cal[1]:.3f}).")
            lines.append("")
        lines += ["| kernel | dispatches | fetch MB / launch | write MB / launch | traffic MB | "
                  "algorithmic MB |", "|---|---|---|---|---|---|"]
        for k in sorted(set(f) & set(wr)):
            if k == "k_copy_probe":
                continue
            fm = np.median(f[k]) * 1024 * (cal[0] if cal else 1) / 1e6
            wm = np.median(wr[k]) * 1024 * (cal[1] if cal else 1) / 1e6
            kid = KERNEL_IDS.get(k)
            alg = f"{table[kid]['bytes_per_launch'] / 1e6:.1f}" if kid in table else "—"
            lines.append(f"| `{k}` | {len(f[k])} | {fm:.1f} | {wm:.1f} | {fm + wm:.1f} | {alg} |")
    open(args.out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
