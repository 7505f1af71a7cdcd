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
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        print(stats_text(args.dir, args.top))
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
