#!/usr/bin/env python3
"""Summaries of rocprofv3 output (kernel statistics and PMC counters).

    python -m hisvsim_amd.profile_summary stats DIR
        per-kernel calls / total / average duration from the
        ``*_kernel_stats.csv`` under DIR
    python -m hisvsim_amd.profile_summary counter DIR COUNTER [KERNEL]
        per-dispatch values of one PMC counter from ``*_counter_collection.csv``
    python -m hisvsim_amd.profile_summary traffic --fetch D1 --write D2
        --cal-fetch D3 --cal-write D4 --kernel K --cal-kernel C --cal-bytes B
        HBM bytes per dispatch of kernel K: its FETCH_SIZE and WRITE_SIZE (KiB
        per dispatch) scaled by bytes / counter-KiB measured on kernel C, a
        copy kernel that reads and writes B bytes each
    python -m hisvsim_amd.profile_summary steps DIR [--n 33 --depth 40 --k 13 --beam 16,8]
        durations of the wide BF16-split part passes in a ``--kernel-trace``
        run of the bench default, grouped by the 5-qubit fused steps of each
        part (the bench's chain parts, regenerated here, in dispatch order)

The CSV layouts are read by column name, so extra columns are ignored.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import statistics
import sys
from typing import Dict, List


def _csvs(d: str, suffix: str) -> List[str]:
    return sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))


def short_name(name: str) -> str:
    """Kernel name without the argument list (template arguments kept)."""
    name = name.replace("(anonymous namespace)::", "")
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return name[:i]
    return name


def kernel_stats(d: str) -> List[dict]:
    rows = []
    for path in _csvs(d, "kernel_stats.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                rows.append({"name": short_name(r["Name"]), "calls": int(r["Calls"]),
                             "total_ms": float(r["TotalDurationNs"]) / 1e6,
                             "avg_ms": float(r["AverageNs"]) / 1e6, "pct": float(r["Percentage"])})
    return sorted(rows, key=lambda r: -r["total_ms"])


def counter_values(d: str, counter: str, kernel: str = "") -> List[float]:
    vals: Dict[str, float] = {}
    for path in _csvs(d, "counter_collection.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter or kernel not in r["Kernel_Name"]:
                    continue
                key = f"{path}:{r['Dispatch_Id']}"
                vals[key] = vals.get(key, 0.0) + float(r["Counter_Value"])
    return list(vals.values())


def stats_text(d: str, top: int = 12) -> str:
    rows = kernel_stats(d)
    out = [f"{'kernel':70s} {'calls':>6s} {'total ms':>10s} {'avg ms':>9s} {'%':>6s}"]
    for r in rows[:top]:
        out.append(f"{r['name'][:70]:70s} {r['calls']:6d} {r['total_ms']:10.3f} {r['avg_ms']:9.3f} {r['pct']:6.2f}")
    return "\n".join(out)


def traffic(fetch_dir, write_dir, cal_fetch_dir, cal_write_dir, kernel, cal_kernel, cal_bytes) -> dict:
    f = counter_values(fetch_dir, "FETCH_SIZE", kernel)
    w = counter_values(write_dir, "WRITE_SIZE", kernel)
    cf = counter_values(cal_fetch_dir, "FETCH_SIZE", cal_kernel)
    cw = counter_values(cal_write_dir, "WRITE_SIZE", cal_kernel)
    if not (f and w and cf and cw):
        raise SystemExit("missing counter data")
    fk, wk = statistics.median(f), statistics.median(w)
    cfk, cwk = statistics.median(cf), statistics.median(cw)
    ff, wf = cal_bytes / cfk, cal_bytes / cwk   # bytes per counter KiB
    return {"kernel": kernel, "dispatches": [len(f), len(w)], "fetch_kib": fk, "write_kib": wk,
            "cal_kernel": cal_kernel, "cal_bytes_each_way": cal_bytes, "cal_fetch_kib": cfk,
            "cal_write_kib": cwk, "fetch_factor": ff, "write_factor": wf,
            "fetch_bytes": fk * ff, "write_bytes": wk * wf, "bytes_per_launch": fk * ff + wk * wf}


def trace_dispatches(d: str, kernel: str) -> List[tuple]:
    """(duration ms, grid size, kernel name) per dispatch of the kernels
    whose name contains KERNEL, in start order, from the
    ``*_kernel_trace.csv`` under DIR."""
    rows = []
    for path in _csvs(d, "kernel_trace.csv"):
        with open(path) as f:
            for r in csv.DictReader(f):
                if kernel in r["Kernel_Name"]:
                    ms = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
                    grid = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
                    rows.append((int(r["Start_Timestamp"]), ms, grid, short_name(r["Kernel_Name"])))
    return [row[1:] for row in sorted(rows)]


def trace_durations(d: str, kernel: str, min_ms: float = 0.0) -> List[float]:
    """Per-dispatch durations (ms) of a kernel, in start order."""
    return [row[0] for row in trace_dispatches(d, kernel) if row[0] >= min_ms]


def steps_table(d: str, n: int, depth: int, k: int, beam=(16, 8)) -> dict:
    """Pass time of the first circuit's part passes by the number of 5-qubit
    steps in the part (the trace of one bench run of the chain). Passes over
    only part of the state (the first passes from |0...0>, which skip the
    tiles that are still zero: a smaller grid than the full passes of the
    kernel instantiation) are left out."""
    from .circuit import random_layered
    from . import capi
    from .partition import chain_parts
    parts, _ = chain_parts(random_layered(n, depth, seed=0), k, beam=beam)
    disp = trace_dispatches(d, "k_part_bf")[:len(parts)]
    max_grid: Dict[str, int] = {}
    for _, grid, name in disp:
        max_grid[name] = max(max_grid.get(name, 0), grid)
    by: Dict[int, List[float]] = {}
    partial = 0
    for p, (ms, grid, name) in zip(parts, disp):
        if grid < max_grid[name]:
            partial += 1
            continue
        by.setdefault(len(capi.part_steps(p)), []).append(ms)
    every = trace_dispatches(d, "k_part_bf")
    for _, grid, name in every:
        max_grid[name] = max(max_grid.get(name, 0), grid)
    full = [ms for ms, grid, name in every if grid == max_grid[name]]
    return {"parts": len(parts), "dispatches": len(disp), "partial_passes_left_out": partial,
            "full_passes_in_trace": len(full), "full_pass_mean_ms": statistics.mean(full) if full else None,
            "by_steps": {s: {"count": len(v), "mean_ms": statistics.mean(v), "min_ms": min(v), "max_ms": max(v)}
                         for s, v in sorted(by.items())}}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("dir")
    a.add_argument("--top", type=int, default=12)
    b = sub.add_parser("counter")
    b.add_argument("dir")
    b.add_argument("counter")
    b.add_argument("kernel", nargs="?", default="")
    c = sub.add_parser("traffic")
    for k in ("fetch", "write", "cal_fetch", "cal_write", "kernel", "cal_kernel"):
        c.add_argument("--" + k.replace("_", "-"), required=True)
    c.add_argument("--cal-bytes", type=float, required=True)
    e = sub.add_parser("steps")
    e.add_argument("dir")
    e.add_argument("--n", type=int, default=33)
    e.add_argument("--depth", type=int, default=40)
    e.add_argument("--k", type=int, default=13)
    e.add_argument("--beam", default="16,8", help="the bench's chain planner beam (width,branch)")
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        print(stats_text(args.dir, args.top))
    elif args.cmd == "steps":
        beam = tuple(int(x) for x in args.beam.split(","))
        print(json.dumps(steps_table(args.dir, args.n, args.depth, args.k, beam), indent=1))
    elif args.cmd == "counter":
        v = counter_values(args.dir, args.counter, args.kernel)
        print(json.dumps({"dispatches": len(v), "median": statistics.median(v) if v else None,
                          "min": min(v) if v else None, "max": max(v) if v else None}))
    else:
        print(json.dumps(traffic(args.fetch, args.write, args.cal_fetch, args.cal_write, args.kernel,
                                 args.cal_kernel, args.cal_bytes), indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
