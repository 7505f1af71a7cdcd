#!/usr/bin/env python3
"""Benchmark: time-to-solution of the 33-qubit depth-40 random layered circuit
on one MI355X, with HBM GB/s of the dominant kernel and the per-gate path
beside it.

    python bench.py [--steps K] [--warmup W] [--n 33] [--depth 40]
                    [--mode fusedbf16|fusedbf16x3|part|fused64|fused32|pergate|permute] [--k K]
                    [--layout chain|fixed] [--beam W,B] [--dtype complex128|complex64]
                    [--plan fixed|auto] [--tune KEY=VALUE ...]

One step = one full simulation of the circuit from |0...0> (state reset
included) with the chosen path; the state (2^n amplitudes) stays in HBM. The
default path is the fused-MFMA one: parts of up to 13 qubits, each cut into
steps of 5 qubits fused into FP64 unitaries and applied in one pass with
BF16-split products accumulated in FP32, the 5 lowest slots moving with the
parts (two in-place swap passes restore the standard layout at the end). The
per-gate path is timed on a bounded sample of the same circuit (its first
layers) and its full time-to-solution is extrapolated per gate. Accuracy of
each path is checked against the FP64 CPU oracle on the same generator at a
size the oracle finishes in seconds, and a fused path also against the FP64
part path at the full size. ``--plan auto`` lets the learned cost model
choose the path, width and partition among its candidate plans. One JSON
line is printed last.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from hisvsim_amd import capi  # noqa: E402
from hisvsim_amd.circuit import gate_amplitudes, random_layered  # noqa: E402
from hisvsim_amd.partition import chain_parts, closure_parts, dfs_parts, greedy_parts  # noqa: E402

METRIC = "33q circuit time-to-solution + HBM GB/s; fused-MFMA vs per-gate at ≤1e-6 infid"
HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK = 78.6e12
FP32_MFMA_PEAK = 157.3e12


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


LOW_SLOTS = 5  # qubits the device adds to every LDS-staged part (libhsv default)


def make_parts(gates, mode: str, k: int, partitioner: str = "best", layout: str = "fixed", beam=(16, 8)):
    """Acyclic parts: DAG-aware greedy, depth-first order, qubit-set
    closure, or whichever gives the fewest parts ("best"). For the register
    part kernel, k counts the low qubits the device adds to every part
    anyway, so the partitioner starts each part with them. With the moving
    layout (wide BF16-split parts only) the parts come from chain_parts and
    the low slots follow them (beam: width and branching of the planner's
    search over low sets). Returns (parts, partitioner, lows or None)."""
    if mode in ("fusedbf16", "fusedbf16x3") and k >= 8 and layout == "chain":
        parts, lows = chain_parts(gates, k, LOW_SLOTS, beam=beam)
        return parts, "chain", lows
    # the register kernel and the wide (k >= 8) BF16-split kernel load whole
    # parts as tiles, so their parts start with the low qubits for long runs
    always = range(LOW_SLOTS) if mode in ("part", "fusedbf16x3") or (mode == "fusedbf16" and k >= 8) else ()
    cands = {"greedy": lambda: greedy_parts(gates, k, always=always),
             "dfs": lambda: dfs_parts(gates, k, always=always),
             "closure": lambda: closure_parts(gates, k, always=always)}
    if partitioner != "best":
        return cands[partitioner](), partitioner, None
    parts, name = min(((f(), name) for name, f in cands.items()), key=lambda x: len(x[0]))
    return parts, name, None


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def measured_traffic(kernel: str, n: int, mode: str, k: int):
    """HBM bytes per launch of the dominant kernel for this workload, from
    the committed PMC summary (FETCH_SIZE + WRITE_SIZE per dispatch, each
    scaled by the counter's reading on a copy kernel of known size), or None
    when that workload was not profiled."""
    try:
        with open(TRAFFIC_FILE) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    row = table.get(f"{kernel}|{mode}|n{n}|k{k}")
    return None if row is None else row.get("bytes_per_launch")


FUSED_MODE = {"fused64": "f64", "fused32": "f32split", "fusedbf16": "bf16split", "fusedbf16x3": "bf16x3"}
BF16_MFMA_PEAK = 2.5e15


def exchange_permutation(n: int, m: int = 3):
    """The qubit permutation of one exchange step (svsim-mpi.hpp:139-149, a
    stable partition of the slots): qubits 9..9+m-1 move to the top m slots
    and every other qubit keeps its relative order, so the top m qubits
    move down by m. Slots below 9 are untouched, which keeps 512-amplitude
    runs whole."""
    mid = list(range(9, 9 + m))
    new_slot = [0] * n
    # qubits in ``mid`` move to the top slots, the others fill the rest in order
    rest = [q for q in range(n) if q not in mid]
    for slot, q in enumerate(rest + mid):
        new_slot[q] = slot
    return new_slot


def chain_full_passes(parts, lows, n: int):
    """Which chain passes run over the whole state. Starting from |0...0>,
    hsv_apply_fused_chain skips the tiles with a 1 in a slot whose qubit no
    earlier part has touched; a pass is full when every slot outside its
    tile holds a touched qubit. The tile is part_apply_bf's: the slots of
    the part's qubits (identities left out) and its lows, padded with the
    lowest free slots to at least 12 (13 when slots 0-4 are not all in it).
    The slot layout follows the chain's own rule: each next low not already
    in slots 0-4 swaps slots with the last current low the next part does
    not list."""
    phys = list(range(n))
    touched, full = set(), []
    for p, (part, low) in enumerate(zip(parts, lows)):
        q = {x for g in part if g.name.upper() != "I" for x in g.qubits()}
        pmask = {phys[x] for x in q | set(low)}
        kt = max(12, min(14, len(pmask | set(range(min(n, 5))))))
        tmask = set(pmask)
        for b in range(n):
            if len(tmask) >= kt:
                break
            tmask.add(b)
        live = {phys[x] for x in touched}
        full.append(all(b in live for b in range(n) if b not in tmask))
        touched |= q
        if p + 1 < len(parts):
            want = set(lows[p + 1])
            spare = [x for x in low if x not in want]
            for x in lows[p + 1]:
                if phys[x] < 5:
                    continue
                o = spare.pop()
                phys[x], phys[o] = phys[o], phys[x]
    return full


def run_path(s: capi.State, mode: str, gates, parts, lows=None) -> None:
    if mode == "permute":
        s.permute_qubits(parts)
        return
    s.reset()
    if mode == "pergate":
        s.apply_gates(gates)
    elif mode == "part":
        for p in parts:
            s.apply_part(p, 0)
    else:
        fm = FUSED_MODE[mode]
        if lows is not None:
            # the split-precision products drift the norm by ~1e-5: the chain
            # renormalises within its layout-restoring swap passes
            s.apply_fused_chain(parts, lows, fm, renormalize=True)
        else:
            for p in parts:
                s.apply_fused(p, fm)
            if fm != "f64":
                s.renormalize()


def verify_small(mode: str, k: int, depth: int, n: int = 20, seed: int = 0, partitioner: str = "best",
                 layout: str = "fixed", dtype: str = "complex128") -> dict:
    """Infidelity and max |d amp| of the chosen path against the FP64 oracle."""
    from oracle.oracle import Oracle
    if mode == "permute":
        rng = np.random.default_rng(seed)
        a = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
        perm = exchange_permutation(n, k)
        want = Oracle().permute(a, perm)
        with capi.State(n) as s:
            s.write(a)
            run_path(s, mode, [], perm)
            got = s.read()
        return {"n": n, "infidelity": None, "max_abs_err": float(np.abs(got - want).max()),
                "bit_exact": bool(np.array_equal(got, want))}
    gates = random_layered(n, depth, seed=seed)
    parts, _, lows = make_parts(gates, mode, k, partitioner, layout) if mode != "pergate" else ([], None, None)
    want = np.zeros(2 ** n, complex)
    want[0] = 1
    o = Oracle()
    o.apply_gates(want, gates)
    with capi.State(n, dtype) as s:
        run_path(s, mode, gates, parts, lows)
        got = s.read()
    # normalised overlap: complex64 storage leaves |got|^2 a few 1e-8 off 1
    infid = max(0.0, 1.0 - abs(np.vdot(want, got)) ** 2 / (np.vdot(want, want).real * np.vdot(got, got).real))
    return {"n": n, "depth": depth, "infidelity": infid, "max_abs_err": float(np.abs(got - want).max())}


def _runs(mask: int) -> int:
    """Contiguous runs of set bits (the reference's loop() nests one level
    per run and indexes past its tables beyond 10, loop.hpp:922-960)."""
    return bin(mask & ~(mask << 1)).count("1")


def cpu_baseline(n_full: int, depth_full: int, k: int, seconds: float = 12.0) -> dict:
    """The reference build (oracle/_ref) timed on the host cores on bounded
    samples of the same circuit generator, each scaled to the full circuit:

    * its partitioned hot path, gather_and_execute_on_multi
      (execute.hpp:318-470) with opt_slots 1 on parts of <= 10 qubits from
      the qubit-set closure partitioner, timed on the first parts of the
      circuit at 28 qubits and scaled per part pass and amplitude to the
      full circuit's parts (this is `value`);
    * its per-gate path, StateVector::apply (state_vector.hpp:402-510), on
      the first layers at 26 qubits, scaled by gate-amplitude updates."""
    from oracle.oracle import RefSim
    if not RefSim.available():
        return {"value": None, "unit": "s", "cores": None, "kind": "reference",
                "sample": "oracle/_ref/libref.so not present"}
    cores = min(len(os.sched_getaffinity(0)), 64)
    ref = RefSim()
    ref.set_threads(cores)
    # per-gate path
    n = min(n_full, 26)
    gates = random_layered(n, 1, seed=0)
    t = ref.time_from_zero(n, gates=gates)
    layers = max(1, min(depth_full, int(seconds / max(t, 1e-3))))
    gates = random_layered(n, layers, seed=0)
    upd = sum(gate_amplitudes(g, n) for g in gates)
    t = ref.time_from_zero(n, gates=gates)
    rate = upd / t
    full_gates = random_layered(n_full, depth_full, seed=0)
    full_upd = sum(gate_amplitudes(g, n_full) for g in full_gates)
    per_gate = {"value": full_upd / rate, "unit": "s", "n_sample": n, "layers": layers, "gates": len(gates),
                "sample_s": t, "gate_amp_updates_per_s": rate}
    # partitioned path: parts the reference's loop() accepts (<= 10 runs
    # inside and outside the part, qubits 0-3 added by opt_slots)
    kp, np_ = 10, min(n_full, 28)
    all_np = (1 << np_) - 1

    def accepted(part):
        m = 0xF
        for g in part:
            for q in g.qubits():
                m |= 1 << q
        return _runs(m) <= 10 and _runs(all_np & ~m) <= 10
    parts = [p for p in closure_parts(random_layered(np_, depth_full, seed=0), kp) if accepted(p)]
    t1 = ref.time_from_zero(np_, parts=parts[:1])
    m = max(1, min(len(parts), int(seconds / max(t1, 1e-3))))
    # m parts spread evenly over the circuit (the timing does not depend on
    # the amplitudes, only on each part's qubits and gates)
    sample = [parts[i] for i in sorted(set(np.linspace(0, len(parts) - 1, m).astype(int)))]
    m = len(sample)
    tp = ref.time_from_zero(np_, parts=sample)
    per_pass_amp = tp / (m * 2 ** np_)
    full_parts = len(closure_parts(full_gates, kp))
    value = per_pass_amp * full_parts * 2 ** n_full
    return {"value": value, "unit": "s", "cores": cores, "kind": "reference",
            "sample": f"reference partitioned path (gather_and_execute_on_multi, opt_slots 1, OpenMP, {cores} "
                      f"threads) on {m} closure parts (<= {kp} qubits, spread over the circuit) of the same generator at "
                      f"{np_} qubits ({tp:.1f} s, {per_pass_amp * 1e9:.2f} ns per amplitude and part pass), "
                      f"extrapolated to the {full_parts} parts of {n_full}q depth {depth_full}; per-gate path "
                      f"(StateVector::apply) beside it: {n}-qubit {layers}-layer prefix ({t:.1f} s), "
                      f"extrapolated by gate-amplitude updates to {per_gate['value']:.0f} s",
            "partitioned": {"value": value, "n_sample": np_, "k": kp, "parts_timed": m, "sample_s": tp,
                            "s_per_amp_pass": per_pass_amp, "full_parts": full_parts},
            "per_gate": per_gate}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=0, help="qubits (default 33; for --mode permute the exchange "
                                                      "runs in place from 32 qubits on, out of place below)")
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--mode", default="fusedbf16",
                    choices=["part", "fused64", "fused32", "fusedbf16", "fusedbf16x3", "pergate", "permute"])
    ap.add_argument("--k", type=int, default=0, help="part width (default: per mode; permute: qubits exchanged)")
    ap.add_argument("--partitioner", default="best", choices=["best", "greedy", "dfs", "closure"],
                    help="acyclic partitioner (best: the one with fewer parts)")
    ap.add_argument("--pergate-layers", type=int, default=1,
                    help="layers of the per-gate timing sample (0: no sample)")
    ap.add_argument("--dtype", default="complex128", choices=["complex128", "complex64"],
                    help="state storage (complex64: the 34-qubit, 137 GB configuration with --n 34)")
    ap.add_argument("--beam", default="16,8", metavar="WIDTH,BRANCH",
                    help="chain planner: beam search over the low sets (1,1 = greedy)")
    ap.add_argument("--layout", default="chain", choices=["chain", "fixed"],
                    help="wide BF16-split parts (k >= 8): moving low-slot layout (chain) or qubits 0-4 "
                         "in every part (fixed)")
    ap.add_argument("--plan", default="fixed", choices=["fixed", "auto"],
                    help="auto: the learned cost model picks mode, width and partition among its candidate "
                         "plans (greedy chain planner), instead of --mode/--k/--partitioner/--beam")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="hsv_tune setting for the timed state (repeatable), e.g. part_tiles_per_wg=2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--no-full-verify", action="store_true",
                    help="skip the full-size infidelity check of a fused mode against the FP64 part path")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): sharding the state over several "
              f"GPUs is a design, not built or run here.", file=sys.stderr)
        return 2
    k = args.k or {"part": 12, "fused64": 5, "fused32": 5, "fusedbf16": 13, "fusedbf16x3": 13, "pergate": 1,
                   "permute": 3}[args.mode]
    n, depth = args.n or 33, args.depth
    gates = random_layered(n, depth, seed=0)
    partitioner_used, lows, plan_s = None, None, None
    beam = tuple(int(x) for x in args.beam.split(","))
    if args.plan == "auto" and args.mode not in ("permute", "pergate"):
        from hisvsim_amd import costmodel
        tp = time.perf_counter()
        model = costmodel.MLP.load()
        try:
            best, cands = costmodel.choose_plan(model, gates, n)
        finally:
            model.close()
        plan_s = time.perf_counter() - tp
        args.mode = {"part": "part", "fused64": "fused64", "fusedbf16": "fusedbf16", "chain": "fusedbf16"}[best["mode"]]
        k, parts, lows = best["k"], best["parts"], best.get("lows")
        partitioner_used = f"cost model pick among {len(cands)} plans: {costmodel.plan_name(best)}"
        beam = (1, 1)
        log(f"[bench] {partitioner_used} (predicted {best['predicted_ms']:.0f} ms), planning {plan_s:.2f} s")
    elif args.mode == "permute":
        parts = exchange_permutation(n, k)
    elif args.mode == "pergate":
        parts = []
    else:
        tp = time.perf_counter()
        parts, partitioner_used, lows = make_parts(gates, args.mode, k, args.partitioner, args.layout, beam)
        plan_s = time.perf_counter() - tp
        log(f"[bench] planning ({partitioner_used}) took {plan_s:.2f} s on the host")
    S = 16 if args.dtype == "complex128" else 8
    total_upd = sum(gate_amplitudes(g, n) for g in gates)
    log(f"[bench] n={n} depth={depth} gates={len(gates)} mode={args.mode} k={k} parts={len(parts)}")

    tune = [(kv.split("=")[0], int(kv.split("=")[1])) for kv in args.tune]
    with capi.State(n, args.dtype) as s:
        for key, value in tune:
            s.tune(key, value)
        for w in range(args.warmup):
            run_path(s, args.mode, gates, parts, lows)
            s.synchronize()
            log(f"[bench] warmup {w} done")
        s.set_timing(True)
        s.kernel_times()
        s.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            run_path(s, args.mode, gates, parts, lows)
            log(f"[bench] step {i} queued")
        s.synchronize()
        t1 = time.perf_counter()
        times = s.kernel_times()
        s.set_timing(False)
        tts = (t1 - t0) / args.steps
        # per-launch durations: one per timed call (pass) in each step
        per_step = len(times) // max(1, args.steps)
        launches = times[: per_step * args.steps].reshape(args.steps, per_step)
        restore_ms, partial = None, None
        if lows is not None:
            if per_step not in (len(parts), len(parts) + 1):
                raise RuntimeError(f"{per_step} timed launches per step for {len(parts)} chain parts")
            if per_step == len(parts) + 1:
                # the moving layout ends with one timed restore (a tile pass,
                # or the two swap passes) back to the standard layout:
                # reported on its own, not in the part average
                restore_ms = float(launches[:, -1].mean())
                launches = launches[:, :-1]
            # from |0...0> the chain's first passes skip the tiles that are
            # still zero: they move less than the whole state, so the
            # per-pass rate below counts the full passes only
            full = chain_full_passes(parts, lows, n)
            partial = {"passes": int(len(full) - sum(full)),
                       "ms": float(launches[:, [i for i, f in enumerate(full) if not f]].sum(axis=1).mean())}
            launches = launches[:, [i for i, f in enumerate(full) if f]]
        kern_ms = float(launches.mean())
        if args.mode == "pergate":
            # the dominant kernel is the dense U3 sweep over the whole state;
            # CZ sweeps move a quarter of it and would skew a per-launch mean
            dense = np.array([g.name == "U3" for g in gates])
            kern_ms = float(launches[:, dense].mean())
        passes = per_step
        log(f"[bench] TTS {tts:.3f} s, {passes} passes/step, mean full pass {kern_ms:.3f} ms"
            + (f", {partial['passes']} partial passes from |0...0> in {partial['ms']:.1f} ms" if partial else ""))

        full_verify = None
        if args.mode in FUSED_MODE and not args.no_full_verify and args.dtype == "complex128":
            # the FP64 part path on a second full-size state (two
            # 2^33 complex128 states fit in HBM together): infidelity of the
            # timed result at the benchmark's own size
            ref_parts, _, _ = make_parts(gates, "part", 12, "best")
            with capi.State(n) as ref:
                run_path(ref, "part", gates, ref_parts)
                ref.synchronize()
                ov = ref.inner(s)
                full_verify = {"n": n, "reference": "FP64 part path (k = 12, tile kernel), same circuit",
                               "infidelity": max(0.0, 1.0 - abs(ov) ** 2 / (ref.norm2() * s.norm2())),
                               "max_abs_err": ref.max_abs_diff(s), "norm2": s.norm2()}
            log(f"[bench] full-size check vs FP64 part path: infidelity {full_verify['infidelity']:.3e}, "
                f"max|d| {full_verify['max_abs_err']:.2e}")

        per_gate = None
        if args.mode != "permute" and args.pergate_layers > 0:
            # per-gate sample: first layers of the same circuit
            per_layer = n + n // 2
            sample = gates[: per_layer * args.pergate_layers]
            s.reset()
            s.synchronize()
            s.set_timing(True)
            s.kernel_times()
            s.apply_gates(sample)
            s.synchronize()
            gt = s.kernel_times()
            s.set_timing(False)
            g_bytes = np.array([2 * S * gate_amplitudes(g, n) for g in sample], float)
            g_ms = gt[: len(sample)]
            dense = np.array([g.name == "U3" for g in sample])
            pergate_tts = float(sum((2 * S * gate_amplitudes(g, n)) for g in gates) /
                                (g_bytes.sum() / (g_ms.sum() * 1e-3)))
            pergate_gbs_dense = float(g_bytes[dense].sum() / (g_ms[dense].sum() * 1e-3) / 1e9)
            pergate_gbs = float(g_bytes.sum() / (g_ms.sum() * 1e-3) / 1e9)
            log(f"[bench] per-gate sample {len(sample)} gates: {pergate_gbs:.0f} GB/s overall, "
                f"{pergate_gbs_dense:.0f} GB/s dense U3; extrapolated per-gate TTS {pergate_tts:.2f} s")
            per_gate = {"sample_gates": len(sample), "gbs": pergate_gbs, "gbs_dense_u3": pergate_gbs_dense,
                        "tts_extrapolated_s": pergate_tts, "passes": len(gates),
                        "speedup_of_mode": pergate_tts / tts}

    if args.mode in ("part", "pergate", "permute", "fusedbf16", "fusedbf16x3"):
        # a BF16-split fused pass executes 3 x 8 x 2^w flops per amplitude for
        # each fused block of width w (one block of k <= 7 qubits, or steps of
        # 5 qubits inside a wide part) against 32 B of HBM traffic: at most
        # ~100 flop/B per block, under the 312 flop/B ridge of 2.5 PFLOP/s
        # over 8 TB/s, so the pass is HBM-bound too
        unit_bytes = 2 * S * 2 ** n
        kernel = {"part": "k_part_reg", "pergate": "k_gate", "permute": "k_permute",
                  "fusedbf16": "k_part_bf" if k >= 8 else "k_fused_bf", "fusedbf16x3": "k_part_bf"}[args.mode]
        launches_per_call = 1
        if args.mode == "part":
            # the tile kernel takes every part that fits one tile with slots
            # 0-4 (12 qubits in complex128, 13 in complex64)
            kt = 12 if args.dtype == "complex128" else 13
            masks = [sum({1 << q for g in p for q in (*g.targets, *g.controls)}) for p in parts]
            if all(bin(m | (2 ** LOW_SLOTS - 1)).count("1") <= kt for m in masks):
                kernel = "k_part_tile"
        if args.mode == "permute" and n >= 32:
            # in place: the permutation as two involutions, one swap-tile
            # pass each (one when the permutation is itself an involution)
            kernel = "k_swap_tiles"
            involution = all(parts[parts[q]] == q for q in range(n))
            launches_per_call = 1 if involution else 2
            kern_ms /= launches_per_call
        achieved = unit_bytes / (kern_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kernel, n, args.mode, k),
                "kernel": kernel,
                "per_launch": f"{unit_bytes} B (2 x {S} B x 2^{n}: read + write of the state) / {kern_ms:.3f} ms"}
        if args.mode in ("fusedbf16", "fusedbf16x3"):
            # blocks per pass: 1 for k <= 7; for wide parts the 5-qubit steps the
            # host cuts (hsv_part_steps); 3 products per real product with the
            # two-piece split, 6 with three pieces
            timed = ([p for p, f in zip(parts, chain_full_passes(parts, lows, n)) if f]
                     if lows is not None else parts)   # the passes kern_ms averages
            wide = k >= 8 or args.mode == "fusedbf16x3"
            blocks = ([len(capi.part_steps(p)) for p in timed] if wide else [1] * len(timed))
            width = 5 if wide else max(5, min(7, k))
            per_pass = float(np.mean(blocks)) * 8 * 2 ** width * 2 ** n
            roof["fused_blocks_per_pass"] = float(np.mean(blocks))
            roof["mfma_tflops_algorithmic"] = per_pass / (kern_ms * 1e-3) / 1e12
            products = 6 if args.mode == "fusedbf16x3" else 3
            roof["mfma_frac_executed"] = products * per_pass / (kern_ms * 1e-3) / BF16_MFMA_PEAK
    else:
        kk = max(4, min(7, k))
        flops = 8 * 2 ** kk * 2 ** n
        peak = FP64_MFMA_PEAK if args.mode == "fused64" else FP32_MFMA_PEAK
        achieved = flops / (kern_ms * 1e-3)
        roof = {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": measured_traffic("k_fused", n, args.mode, k),
                "kernel": "k_fused",
                "per_launch": f"8 x 2^k x 2^{n} algorithmic flops (k={kk}) / {kern_ms:.3f} ms",
                "hbm_gbs": 2 * S * 2 ** n / (kern_ms * 1e-3) / 1e9}

    verify = None if args.no_verify else verify_small(args.mode, k, depth, partitioner=args.partitioner,
                                                      layout=args.layout, dtype=args.dtype)
    if verify:
        log(f"[bench] verify vs FP64 oracle at n={verify['n']}: infidelity {verify['infidelity']}, "
            f"max|d| {verify['max_abs_err']:.2e}")
    cpu = None if args.no_cpu or args.mode == "permute" else cpu_baseline(n, depth, k)
    if cpu:
        log(f"[bench] cpu baseline: {cpu}")

    out = {
        "metric": METRIC, "value": tts, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": tts * 1e3, "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": {"fused32": "f32split", "fusedbf16": "bf16split (fp32 accumulate)",
                  "fusedbf16x3": "bf16x3split (three-piece split, fp32-class products, fp32 accumulate)"}.get(
                      args.mode, "f64"),
        "data": "synthetic: seeded random layered circuit (U3 angles ~ U[0,2pi), CZ on random matchings), |0..0> input",
        "config": ({"workload": f"RL{n} random layered circuit, {n} qubits, depth {depth}, {args.dtype} state "
                                f"({2 ** n * S / 1e9:.1f} GB)", "n": n, "depth": depth, "gates": len(gates),
                    "mode": args.mode, "k": k, "passes": passes, "partitioner": (partitioner_used if lows is None or args.plan == "auto"
                                                                      else f"chain (beam {beam[0]}x{beam[1]})"),
                    "plan": args.plan,
                    # host planning runs once per circuit before the timed
                    # steps, as the reference reads a precomputed parts file:
                    # `value` excludes it
                    "plan_s": plan_s, "plan_host_threads": len(os.sched_getaffinity(0)),
                    **({"layout": "chain (moving low slots, one restoring permutation pass)",
                        "layout_restore_ms": restore_ms, "partial_passes_from_zero": partial}
                       if lows is not None else {}),
                    **({"tune": dict(tune)} if tune else {})}
                   if args.mode != "permute" else
                   {"workload": f"exchange permutation of a {n}-qubit {args.dtype} state ({2 ** n * S / 1e9:.1f} GB): "
                                f"qubits 9..{8 + k} to the top slots", "n": n, "mode": "permute", "k": k,
                    "passes": passes * (launches_per_call if args.mode == "permute" else 1),
                    "form": ("in place (k_swap_tiles involution passes)" if n >= 32 else
                             "out of place (k_permute into a second buffer)")}),
        "hbm_gbs": 2 * S * 2 ** n / (kern_ms * 1e-3) / 1e9,
        "gate_amp_updates_per_s": total_upd / tts,
        "per_gate": per_gate,
        "roofline": roof,
        "verify": verify,
        "verify_full_size": full_verify,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
