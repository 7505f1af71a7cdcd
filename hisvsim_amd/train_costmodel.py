#!/usr/bin/env python3
"""Trains and evaluates the cost model on the GPU.

    python -m hisvsim_amd.train_costmodel [--circuits 2000] [--out DIR]

1. Data: random layered circuits (U3 + random-matching CZ, the benchmark
   generator) with n ~ U{20..30} qubits and depth ~ U{4..10}, and with
   --extra E, E more circuits of controlled-phase families (QFT prefixes
   and layers of H with controlled phases on random pairs), whose parts
   fuse into many more 5-qubit steps than a layered circuit's. Each runs under
   two plans drawn from the candidate set (costmodel.candidate_specs: fused
   FP64 / BF16-split unitaries, wide BF16-split parts, the moving-layout
   chain, the register part kernel), from a state taken off |0...0> (so
   every pass covers the whole state); every part's kernel time is recorded
   with HIP events, and a chain's layout restore on its own.
2. Training: the 3-layer MLP (64 hidden units, FP32) on log(ms per 2^30
   amplitudes), by libhsv's own forward/backward/Adam kernels; 10% held out.
3. Evaluation: 32-qubit random layered circuits of depth ~ U{30..60}. The
   model predicts every candidate plan, picks the cheapest, and the picked
   plan and the fallback (natural-order parts, k = 10) are both run.
   Reported: per-part and per-plan predicted-vs-measured error and the
   end-to-end times.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import capi
from .circuit import Gate, qft, random_layered
from .costmodel import (MLP, compare_choice, PARAMS_FILE, RESTORE_MS_PER_2_30, build_plan, candidate_specs, choose_plan, fallback_plan,
                        ms_of, plan_features, predict_plan_ms, run_plan, target_of)


def log(msg):
    print(msg, flush=True)


def dense_start(s) -> None:
    """Resets the state and makes it dense with one untimed layer of U3
    gates (a product state with every amplitude nonzero). Every pass of a
    plan then runs over the whole state (from |0...0> the chain's first
    passes skip the tiles that are still zero), and on dense data: the
    matrix cores draw more power on dense operands, and a mostly-zero state
    would time the BF16-split passes up to ~25% fast."""
    s.reset()
    s.apply_gates([Gate("U3", (q,), (), (0.7 + 0.05 * q, 0.3, 1.1)) for q in range(s.n)])


def phase_family(n: int, seed: int) -> list:
    """A controlled-phase circuit: a QFT prefix of random length, or layers
    of H on a random half of the qubits and controlled R1 on random pairs."""
    rng = np.random.default_rng(seed)
    if rng.random() < 0.5:
        g = qft(n)
        return g[:int(rng.integers(n, len(g) + 1))]
    out = []
    for _ in range(int(rng.integers(3, 9))):
        out += [Gate("H", (int(q),)) for q in rng.choice(n, n // 2, replace=False)]
        for _ in range(n):
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            out.append(Gate("R1", (a,), (b,), (float(rng.uniform(0, 2 * np.pi)),)))
    return out


def collect(num: int, seed: int, n_range=(20, 30), depth_range=(4, 10), max_seconds: float = 0.0, extra: int = 0):
    """Circuits are drawn first and run in blocks of 100, each block in
    ascending n (one state allocation per n and block). With max_seconds > 0
    collection stops at the first block boundary past that time. extra:
    circuits of the controlled-phase families (phase_family) after the num
    layered ones (depth -1 in the job list)."""
    rng = np.random.default_rng(seed)
    jobs = []
    specs = candidate_specs()
    for c in range(num + extra):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        depth = int(rng.integers(depth_range[0], depth_range[1] + 1)) if c < num else -1
        picks = [int(i) for i in rng.choice(len(specs), size=2, replace=False)]
        jobs.append((n, depth, int(rng.integers(1 << 30)), picks))
    jobs = [j for b in range(0, num + extra, 100) for j in sorted(jobs[b:b + 100], key=lambda j: j[0])]
    X, y, meta, restore = [], [], [], []
    t0 = time.time()
    s = None
    for c, (n, depth, gseed, picks) in enumerate(jobs):
        if max_seconds > 0 and c % 100 == 0 and c and time.time() - t0 > max_seconds:
            log(f"[collect] time budget reached after {c} circuits")
            break
        gates = random_layered(n, depth, seed=gseed) if depth >= 0 else phase_family(n, gseed)
        if s is None or s.n != n:
            if s is not None:
                s.close()
            s = capi.State(n)
            dense_start(s)   # stays dense: every plan is unitary (BF16 plans renormalise)
        for pi in picks:
            plan = build_plan(gates, n, specs[pi])
            times = run_plan(s, plan)
            for f, ms in zip(plan_features(plan, n), times):
                X.append(f)
                y.append(target_of(float(ms), n))
            if plan["mode"] == "chain" and len(times) == len(plan["parts"]) + 1:
                restore.append((n, float(times[-1])))
            meta.append({"n": n, "depth": depth, "mode": plan["mode"], "k": plan["k"],
                         "partitioner": plan["partitioner"], "parts": len(plan["parts"]),
                         "ms": float(times.sum())})
        if (c + 1) % 50 == 0:
            log(f"[collect] {c + 1}/{num + extra} circuits (n={n}), {len(X)} parts, {time.time() - t0:.0f} s")
    if s is not None:
        s.close()
    return np.array(X, np.float32), np.array(y, np.float32), meta, np.array(restore, np.float64).reshape(-1, 2)


def restore_coefficient(restore: np.ndarray) -> float:
    """Median layout-restore time of the chain plans, per 2^30 amplitudes."""
    if not len(restore):
        return RESTORE_MS_PER_2_30
    return float(np.median(restore[:, 1] / 2.0 ** (restore[:, 0] - 30)))


def evaluate(model: MLP, num: int, seed: int, n: int = 32):
    """32-qubit random layered circuits of depth U{30..60}: for each, the
    model's pick, its next-ranked candidates, the fixed default (the
    moving-layout chain at k = 13) and the plain partitioner are all run
    and timed (costmodel.compare_choice)."""
    rng = np.random.default_rng(seed)
    rows = []
    with capi.State(n) as s:
        dense_start(s)
        for c in range(num):
            depth = int(rng.integers(30, 61))
            gates = random_layered(n, depth, seed=int(rng.integers(1 << 30)))
            res = compare_choice(model, gates, n, s)
            res.update({"depth": depth, "gates": len(gates)})
            log(f"[eval] circuit {c} depth {depth}: pick {res['pick']} regret {res['regret']:.3f}, "
                f"gain over default {res['gain_over_default']:.3f}, over plain {res['gain_over_plain']:.2f}, "
                f"plan error {res['plan_rel_err']:.3f}")
            rows.append(res)
    part_err = [e for r in rows for e in r["part_rel_err"]]
    for r in rows:
        del r["part_rel_err"]
    return rows, float(np.median(part_err)), float(np.mean(part_err))


def family_circuits():
    """Circuit families the default plan (the chain at k = 13) is not built
    for, beside two layered ones it is: (name, n, gates)."""
    from .circuit import bernstein_vazirani
    ghz = [Gate("H", (0,))] + [Gate("X", (q + 1,), (q,)) for q in range(29)]
    return [("rl28_depth3", 28, random_layered(28, 3, seed=5)),
            ("rl30_depth6", 30, random_layered(30, 6, seed=6)),
            ("qft28", 28, qft(28)),
            ("bv28", 28, bernstein_vazirani(28, seed=3)[0]),
            ("ghz30", 30, ghz),
            ("rl32_depth20", 32, random_layered(32, 20, seed=9))]


def compare_families(model: MLP) -> list:
    """costmodel.compare_choice on each of family_circuits(), from a dense
    state: the model's pick, its next-ranked candidates, the default and the
    plain partitioner, measured."""
    rows = []
    for name, n, gates in family_circuits():
        with capi.State(n) as s:
            dense_start(s)
            res = compare_choice(model, gates, n, s)
        del res["part_rel_err"]
        res.update({"family": name, "n": n, "gates": len(gates)})
        log(f"[families] {name}: pick {res['pick']} regret {res['regret']:.3f}, "
            f"gain over default {res['gain_over_default']:.3f}, over plain {res['gain_over_plain']:.2f}")
        rows.append(res)
    return rows


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--circuits", type=int, default=2000)
    ap.add_argument("--eval", type=int, default=3)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--out", default="gpurun_out")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--stage", choices=("all", "collect", "fit", "families"), default="all",
                    help="collect: time circuits and save the samples to --data; fit: train and evaluate "
                         "on the samples in --data (several files, comma-separated); all: both; families: "
                         "the committed model's choice measured on other circuit families")
    ap.add_argument("--data", default="")
    ap.add_argument("--max-seconds", type=float, default=0.0, help="time budget of the collection")
    ap.add_argument("--extra", type=int, default=0,
                    help="circuits of the controlled-phase families on top of --circuits layered ones")
    args = ap.parse_args(argv)
    os.makedirs(args.out, exist_ok=True)
    if args.stage == "families":
        model = MLP.load()
        try:
            rows = compare_families(model)
        finally:
            model.close()
        with open(os.path.join(args.out, "costmodel_families.json"), "w") as f:
            json.dump(rows, f, indent=1)
        return 0
    if args.stage == "fit":
        parts = [np.load(p) for p in args.data.split(",") if p]
        X = np.concatenate([d["X"] for d in parts])
        y = np.concatenate([d["y"] for d in parts])
        restore = np.concatenate([d["restore"] for d in parts if "restore" in d.files] or [np.zeros((0, 2))])
        meta = []
        args.circuits = int(sum(int(d["circuits"]) for d in parts))
    else:
        X, y, meta, restore = collect(args.circuits, args.seed, max_seconds=args.max_seconds, extra=args.extra)
        args.circuits = len(meta) // 2
        if args.data:
            np.savez(args.data, X=X, y=y, circuits=args.circuits, restore=restore)
            log(f"[collect] saved {len(X)} samples to {args.data}")
        if args.stage == "collect":
            return 0
    log(f"[train] {len(X)} samples")
    rng = np.random.default_rng(1)
    idx = rng.permutation(len(X))
    nval = max(1, len(X) // 10)
    val, tr = idx[:nval], idx[nval:]
    model = MLP(seed=7)
    model.restore_ms_per_2_30 = restore_coefficient(restore)
    log(f"[train] chain layout restore: {model.restore_ms_per_2_30:.2f} ms per 2^30 amplitudes "
        f"({len(restore)} samples)")
    t0 = time.time()
    loss = model.train(X[tr], y[tr], epochs=args.epochs, batch=512, lr=2e-3)
    train_s = time.time() - t0
    pv = model.predict(X[val])
    val_mse = float(np.mean((pv - y[val]) ** 2))
    val_rel = float(np.median(np.abs(np.exp(pv - y[val]) - 1)))
    log(f"[train] loss {loss:.4f}, held-out MSE {val_mse:.4f}, median relative error {val_rel:.3f}, {train_s:.1f} s")
    rows, med, mean = evaluate(model, args.eval, args.seed + 1)
    report = {"samples": len(X), "circuits": args.circuits, "train_loss": loss, "val_mse_log": val_mse,
              "val_median_rel_err": val_rel, "train_seconds": train_s,
              "eval_part_median_rel_err": med, "eval_part_mean_rel_err": mean,
              "eval_regret_max": max(r["regret"] for r in rows) if rows else None,
              "eval_gain_over_default_min": min(r["gain_over_default"] for r in rows) if rows else None,
              "eval_gain_over_plain_min": min(r["gain_over_plain"] for r in rows) if rows else None,
              "eval_plan_rel_err_max": max(r["plan_rel_err"] for r in rows) if rows else None,
              "eval": rows,
              "collect_plans": meta[:50]}
    with open(os.path.join(args.out, "costmodel_report.json"), "w") as f:
        json.dump(report, f, indent=1)
    model.save(os.path.join(args.out, "costmodel_params.json"),
               meta={"samples": len(X), "val_median_rel_err": val_rel,
                     "restore_ms_per_2_30": model.restore_ms_per_2_30})
    log(json.dumps({k: v for k, v in report.items() if k not in ("eval", "collect_plans")}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
