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
    "k_cell_keys": "cell_keys", "k_cell_scatter": "cell_scatter",
    "k_neighbors": "neighbors", "k_neighbors<true>": "neighbors", "k_neighbors<false>": "neighbors", "k_density<true, true, false>": "density_ecs",
    "k_density<true, false, true>": "density_save", "k_explicit": "explicit",
    "k_ppe_count": "ppe_count", "k_ppe_fill": "ppe_fill", "k_cg_init": "cg_init",
    "k_ppe_fill<true>": "ppe_fill", "k_ppe_fill<false>": "ppe_fill",
    "k_cg_spmv<false>": "cg_spmv", "k_cg_spmv<true>": "cg_spmv", "k_cg_update": "cg_update",
    "k_cg_spmv<false, mps::(anonymous namespace)::Sell<double> >": "cg_spmv",
    "k_cg_spmv<true, mps::(anonymous namespace)::Sell<double> >": "cg_spmv",
    "k_gradient": "gradient", "k_ds": "ds",
    "k_ml_spmv<false, mps::(anonymous namespace)::MatFree>": "cg_spmv32",
    "k_ml_spmv<true, mps::(anonymous namespace)::MatFree>": "cg_spmv32",
    "k_ml_spmv<false, mps::(anonymous namespace)::MatFreeRec>": "cg_spmv32",
    "k_ml_spmv<true, mps::(anonymous namespace)::MatFreeRec>": "cg_spmv32",
    "k_density_explicit": "density_ecs", "k_density_ppe_count": "density_save",
    "k_ml_prec": "ml_prec", "k_ml_prec<0>": "ml_prec", "k_ml_prec<1>": "ml_prec", "k_ml_prec<2>": "ml_prec",
    "k_ml_prec_up": "ml_prec",
    "k_cgf_iter<mps::(anonymous namespace)::MatFree>": "cg_iter32",
    "k_features<unsigned char>": "features", "k_features<unsigned short>": "features",
}


# algorithmic bytes per launch of the kernels the library does not time on
# their own, from the workload's sizes (N rows, Z stored off-diagonals)
FORMULAS = {
    "k_apply_du": lambda N, Z: 81 * N,  # u, w, x, z read and written; type
    "k_explicit_update": lambda N, Z: 129 * N,  # and the packed (x, z), (u, w) pairs written
    "k_pack_state": lambda N, Z: 64 * N,
    "k_writeback": lambda N, Z: 17 * N,  # type, sol read; p written
    "k_ir_correct": lambda N, Z: 20 * N,  # x read and written, d read
    "k_jacobi_inv": lambda N, Z: 12 * N,
    "k_rec_init": lambda N, Z: 24 * N,
    "k_rec_p": lambda N, Z: 32 * N,
    "k_apply_pred": lambda N, Z: 21 * N,
    "k_revert_unless_better": lambda N, Z: 17 * N,
    # b, p, diag, type, x, z; the records written; length, ghost and own
    # record of the operator's row; 2-byte column offsets
    "k_learn_resid32": lambda N, Z: 97 * N + 2 * Z,
    "k_cg_init": lambda N, Z: 12 * Z + 32 * N,
    "k_cg_init<mps::(anonymous namespace)::Sell<double> >": lambda N, Z: 12 * Z + 32 * N,
    # sorted ids and x, z read, the cell-ordered positions written (k_cell_sort
    # shares the library's statistic with it and is not shown in bytes)
    "k_cell_pos": lambda N, Z: 36 * N,
    # x, z, n, type, row length, the records and the row written; column offsets
    "k_features<unsigned char>": lambda N, Z: 381 * N + 2 * Z,
    "k_features<unsigned short>": lambda N, Z: 381 * N + 2 * Z,
}
MLP_FLOP_PER_ROW = 2 * (160 * 256 + 256 * 256 + 256)
DENSE_BF16_PEAK_TFLOPS = 2500.0


def sizes(bench: dict):
    """N and Z of the profiled workload: N from the config, Z from a kernel
    whose accounted bytes it determines (FP64 SpMV 12 Z + 48 N, or the
    feature pass 2 Z + 381 N)."""
    N = bench.get("config", {}).get("particles")
    Z = None
    k = {**bench.get("kernels", {}), **bench.get("kernels_fp64", {})}
    if N and "cg_spmv" in k:
        Z = (k["cg_spmv"]["bytes_per_launch"] - 48 * N) / 12
    elif N and "features" in k:
        Z = (k["features"]["bytes_per_launch"] - 381 * N) / 2
    return N, Z


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
              "launch), from the bench JSON of the same command, or for kernels the library does not",
              "time on their own from the workload's sizes (N rows, Z off-diagonals; see FORMULAS in",
              "openmps_amd/profile_summary.py); GB/s = bytes / mean working duration.  The MLP",
              "inference is FLOP-bound: TFLOP/s against the 2.5 PFLOP/s dense BF16 peak.",
              "", "| kernel | calls | total ms | % time | mean working us | MB / launch | GB/s | % of 8 TB/s |",
              "|---|---|---|---|---|---|---|---|"]
    N, Z = sizes(bench)
    for k in sorted(dur, key=lambda k: -dur[k].sum()):
        a = dur[k]
        w = working(a)
        kid = KERNEL_IDS.get(k)
        mb = gbs = frac = "—"
        bpl = None
        if kid and kid in table and "bytes_per_launch" in table[kid]:
            bpl = table[kid]["bytes_per_launch"]
        elif k in FORMULAS and N and Z is not None:
            bpl = FORMULAS[k](N, Z)
        elif k == "k_copy_probe" and "copy_kernel" in bench:
            bpl = bench["copy_kernel"]["bytes_moved"]
        if bpl is not None:
            mb = f"{bpl / 1e6:.1f}"
            g = bpl / (w.mean() * 1e3)
            gbs, frac = f"{g:.0f}", f"{100 * g / HBM_PEAK_GBS:.1f}"
        elif k == "k_mlp_infer" and N:
            tf = MLP_FLOP_PER_ROW * N / (w.mean() * 1e-6) / 1e12
            mb = "(FLOP)"
            gbs, frac = f"{tf:.0f} TFLOP/s", f"{100 * tf / DENSE_BF16_PEAK_TFLOPS:.1f} of BF16"
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
            alg = "—"
            if kid in table:
                alg = f"{table[kid]['bytes_per_launch'] / 1e6:.1f}"
            elif k in FORMULAS and N and Z is not None:
                alg = f"{FORMULAS[k](N, Z) / 1e6:.1f}"
            lines.append(f"| `{k}` | {len(f[k])} | {fm:.1f} | {wm:.1f} | {fm + wm:.1f} | {alg} |")
    open(args.out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
