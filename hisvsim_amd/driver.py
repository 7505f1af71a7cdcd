#!/usr/bin/env python3
"""QASM driver: parse a circuit, partition it, run it on the GPU, measure.

Counterpart of the reference driver's main flow and part executor
(qasm_assembler_standalone.py:3765-3870 and ``execute_qasm_parts`` :1768-2030,
``execute_qasm`` :1377-1617 for the per-gate path):

    python -m hisvsim_amd.driver --app circuit.qasm [--parts --partfile F]
        [--nat K | --greedy K] [--slots] [--mode part|fused64|fused32|fusedbf16|fusedbf16x3|pergate]
        [--layout fixed|chain [--beam W,B]] [--plan file|auto]
        [--lowering reference|textbook] [--verify auto|none|fp64|inverse]
        [--dot] [--dump] [--shots N]
        [--n T --alloc A --initn T]   (accepted for compatibility; no effect)

Parts come from a parts file (``<order> <node> <part>`` lines, the reference
format), or from this package's natural-order (``--nat K``, the reference's
``*_part_nat`` method) or DAG-aware greedy partitioner (``--greedy K``).
Without ``--parts`` every gate is applied on its own (the per-gate path).
``--slots`` adds the fast low qubits to every part (``set_opt_slots(1)``).
``--mode fusedbf16 --layout chain`` runs the benchmark's path: parts of up
to ``--nat``/``--greedy`` K (default 13) qubits from the moving-layout
planner (``chain_parts``, beam ``--beam``) through hsv_apply_fused_chain.
``--mode fusedbf16x3`` is the three-piece BF16 split (FP32-class products;
parts whose qubits and qubits 0-4 fit 13).
``--plan auto`` lets the learned cost model (``costmodel.choose_plan``)
pick the mode, width and partition among its candidates; without a trained
parameter file (or device for the network) it falls back to the plain
partitioner (natural order, k = 10, part kernel).
``--verify`` reports the run's accuracy (the reference's ``verify_compute``,
:716-1198): ``fp64`` runs the circuit again on the FP64 per-gate path in a
second state and reports the infidelity and max |a - b|; ``inverse`` applies
the inverse circuit to the result on the FP64 per-gate path and checks that
it returns to |0...0>; ``auto`` (the default) does ``fp64`` for the
split-precision modes (FP32-split, BF16-split) and nothing for the FP64
ones. ``--dot`` writes the DAG files next to the input, ``--dump`` the
``<file>_exec`` ordered-gate dump. Measurements in the file are taken at
the end (deferred): a gate on an already measured qubit is an error. One
JSON line reports the timings (µs, like the reference's counters) and, per
measured qubit, P(1) and ``--shots`` samples of the measured register.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from typing import List

import numpy as np

from . import capi, qasm, verify
from .partition import check_parts, greedy_parts, nat_parts

NUM_MEASURE = 10   # shots per measurement in the reference driver (qasm_assembler_standalone.py:29)


def measured_qubits(prog: qasm.QasmProgram) -> List[int]:
    """Qubit ids measured in the file, in order; raises if a gate acts on a
    qubit after it was measured (the deferred-measurement condition)."""
    ids = prog.qubit_ids()
    done: List[int] = []
    for op in prog.ops:
        if op.gate == "measure":
            q = ids[op.qubits[0]]
            if q not in done:
                done.append(q)
        elif any(ids[x] in done for x in op.qubits):
            raise qasm.QasmError(f"line {op.index + 1}: gate on a qubit measured earlier "
                                 f"(mid-circuit measurement is not supported)")
    return done


def plan_parts(prog: qasm.QasmProgram, args) -> List[list]:
    gates = prog.gates(args.lowering)
    if args.layout == "chain":
        from .partition import chain_parts
        beam = tuple(int(x) for x in args.beam.split(","))
        parts, lows = chain_parts(gates, args.nat or args.greedy or 13, 5, beam=beam)
        args.lows = lows
        return parts
    if args.partfile:
        parts = qasm.parts_from_parts_file(prog, args.partfile, args.lowering)
    elif args.nat:
        parts = nat_parts(gates, args.nat)
    elif args.greedy:
        parts = greedy_parts(gates, args.greedy)
    else:
        parts = nat_parts(gates, 10)
    check_parts(gates, parts)
    return parts


def auto_plan(gates, n: int):
    """The cost model's pick among its candidate plans, or the plain
    partitioner when no trained model is at hand."""
    from . import costmodel
    try:
        model = costmodel.MLP.load()
    except (OSError, ValueError, KeyError, capi.HsvError) as e:
        plan = costmodel.fallback_plan(gates, n)
        return plan, {"source": "fallback", "reason": f"no cost model ({e})", "mode": plan["mode"],
                      "k": plan["k"], "partitioner": plan["partitioner"]}
    try:
        best, plans = costmodel.choose_plan(model, gates, n)
    finally:
        model.close()
    return best, {"source": "cost model", "mode": best["mode"], "k": best["k"], "partitioner": best["partitioner"],
                  "predicted_ms": best["predicted_ms"], "candidates": len(plans)}


def run(args) -> dict:
    t0 = time.perf_counter()
    prog = qasm.parse_file(args.app)
    gates = prog.gates(args.lowering)
    meas = measured_qubits(prog)
    t_parse = time.perf_counter() - t0
    if args.dot:
        qasm.write_dot_files(prog)
    if args.dump:
        with open(prog.filename + "_exec", "w") as f:
            f.write(qasm.exec_dump(prog))
    n = prog.num_qubits
    args.lows = None
    plan_info = None
    if args.plan == "auto":
        plan, plan_info = auto_plan(gates, n)
        args.parts = True
        args.mode = {"part": "part", "fused64": "fused64", "fusedbf16": "fusedbf16", "chain": "fusedbf16"}[plan["mode"]]
        args.lows = plan.get("lows")
        parts = plan["parts"]
    else:
        parts = plan_parts(prog, args) if args.parts else []
    dtype = "complex128" if args.dtype == "auto" else args.dtype
    with capi.State(n, dtype) as s:
        s.synchronize()
        t1 = time.perf_counter()
        if not args.parts or args.mode == "pergate":
            s.apply_gates(gates)
        elif args.mode == "part":
            for p in parts:
                s.apply_part(p, 1 if args.slots else 0)
        elif args.lows is not None:
            s.apply_fused_chain(parts, args.lows, "bf16x3" if args.mode == "fusedbf16x3" else "bf16split",
                                renormalize=True)
        else:
            fm = {"fused64": "f64", "fused32": "f32split", "fusedbf16": "bf16split", "fusedbf16x3": "bf16x3"}[args.mode]
            for p in parts:
                s.apply_fused(p, fm)
            if fm != "f64":
                s.renormalize()
        s.synchronize()
        t_sim = time.perf_counter() - t1
        probs = {q: s.probability(q, 1) for q in meas}
        samples = (s.measure(meas, np.random.default_rng(args.seed).random(args.shots))
                   if meas and args.shots else [])
        norm2 = s.norm2()
        mode = args.mode if args.parts else "pergate"
        check = getattr(args, "verify", "auto")
        if check == "auto":
            check = "fp64" if mode in ("fused32", "fusedbf16", "fusedbf16x3") else "none"
        accuracy = None
        if check == "fp64":
            accuracy = verify.against_fp64(s, gates)
        elif check == "inverse":
            accuracy = verify.by_inverse(s, gates)
    return {
        "app": args.app, "qubits": n, "gates": len(gates), "parts": len(parts) if args.parts else len(gates),
        "mode": args.mode if args.parts else "pergate", "lowering": args.lowering, "dtype": dtype,
        "layout": "chain" if args.lows is not None else "fixed", "plan": plan_info,
        "compile_us": int(t_parse * 1e6), "sim_us": int(t_sim * 1e6), "norm2": norm2,
        "measured": meas, "p1": {str(q): p for q, p in probs.items()},
        "samples": [list(map(int, x)) for x in samples], "accuracy": accuracy,
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--app", required=True, help="QASM file")
    ap.add_argument("--parts", action="store_true", help="run partitioned (else per gate)")
    ap.add_argument("--partfile", default="", help="parts file (<order> <node> <part> lines)")
    ap.add_argument("--nat", type=int, default=0, help="natural-order parts of K qubits")
    ap.add_argument("--greedy", type=int, default=0, help="DAG-aware greedy parts of K qubits")
    ap.add_argument("--slots", action="store_true", help="add the fast low qubits to every part")
    ap.add_argument("--mode", default="part",
                    choices=["part", "fused64", "fused32", "fusedbf16", "fusedbf16x3", "pergate"])
    ap.add_argument("--layout", default="fixed", choices=["fixed", "chain"],
                    help="chain: moving low-slot layout (fusedbf16 / fusedbf16x3; parts of --nat/--greedy K, "
                         "default 13)")
    ap.add_argument("--beam", default="16,8", metavar="WIDTH,BRANCH", help="chain planner beam (1,1 = greedy)")
    ap.add_argument("--plan", default="file", choices=["file", "auto"],
                    help="auto: the learned cost model picks mode, width and partition (fallback: nat, k = 10)")
    ap.add_argument("--lowering", default="reference", choices=["reference", "textbook"])
    ap.add_argument("--verify", default="auto", choices=["auto", "none", "fp64", "inverse"],
                    help="accuracy report: fp64 reference run, inverse circuit, or fp64 for split modes")
    ap.add_argument("--dtype", default="auto", choices=["auto", "complex128", "complex64"],
                    help="state precision (auto: complex128)")
    ap.add_argument("--dot", action="store_true", help="write <file>_PART_<i>.dot")
    ap.add_argument("--dump", action="store_true", help="write <file>_exec")
    ap.add_argument("--shots", type=int, default=NUM_MEASURE)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--n", type=int, default=0, help="host threads (no effect on the GPU path)")
    ap.add_argument("--alloc", type=int, default=0, help="allocation scheme (no effect on the GPU path)")
    ap.add_argument("--initn", type=int, default=0, help="init threads (no effect on the GPU path)")
    args = ap.parse_args(argv)
    if args.layout == "chain" and (args.mode not in ("fusedbf16", "fusedbf16x3") or not args.parts):
        ap.error("--layout chain runs partitioned on a BF16-split path: add --parts --mode fusedbf16")
    print(json.dumps(run(args)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
