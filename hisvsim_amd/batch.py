"""Batched small circuits on one GPU (hsv_batch_*): many circuits, or
parameter sets of one circuit, of uneven qubit counts in one hipGraph.

    with CircuitBatch([(n0, gates0), (n1, gates1), ...]) as b:
        b.run()                       # every instance from |0...0>
        amps = b.state(i)
        z = b.expect_z([mask0, mask1, ...])
        b.set_gates([gates0', ...])   # same structure, new parameters: same graph
        b.run()

``python -m hisvsim_amd.batch`` measures the benchmark batch: parameter
sets of a hardware-efficient ansatz (RY/RZ layers and CZ chains) on uneven
qubit counts, with the schedule's HBM traffic against the 8 TB/s peak.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import sys
from typing import List, Sequence, Tuple

import numpy as np

from . import capi
from .circuit import Gate

TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_batch.json")


class CircuitBatch:
    """Instances (n, gates); all start from |0...0>."""

    def __init__(self, instances: Sequence[Tuple[int, Sequence[Gate]]], dtype: str = "complex128"):
        self.instances = [(int(n), list(g)) for n, g in instances]
        self.dtype = dtype
        nq = (ctypes.c_int * len(self.instances))(*[n for n, _ in self.instances])
        offs = [0]
        for _, g in self.instances:
            offs.append(offs[-1] + len(g))
        self._offs = offs
        go = (ctypes.c_int * len(offs))(*offs)
        arr = capi.gate_array([g for _, gs in self.instances for g in gs])
        self._h = ctypes.c_void_p()
        dt = {"complex128": capi.HSV_C128, "complex64": capi.HSV_C64}[dtype]
        capi._check(capi.lib().hsv_batch_create(len(self.instances), nq, go, arr, dt, ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h:
            capi.lib().hsv_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_gates(self, gate_lists: Sequence[Sequence[Gate]]) -> None:
        """New gates with the same types and qubits, instance by instance."""
        flat = [g for gs in gate_lists for g in gs]
        capi._check(capi.lib().hsv_batch_set_gates(self._h, capi.gate_array(flat)))
        self.instances = [(n, list(g)) for (n, _), g in zip(self.instances, gate_lists)]

    def run(self) -> None:
        capi._check(capi.lib().hsv_batch_run(self._h))

    def state(self, i: int) -> np.ndarray:
        n = self.instances[i][0]
        out = np.zeros(2 * 2 ** n)
        capi._check(capi.lib().hsv_batch_read(self._h, i, out.ctypes.data_as(capi._D)))
        return out[0::2] + 1j * out[1::2]

    def expect_z(self, masks: Sequence[int]) -> np.ndarray:
        m = np.array([int(x) for x in masks], dtype=np.uint64)
        out = np.zeros(len(self.instances))
        capi._check(capi.lib().hsv_batch_expect_z(self._h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                                  out.ctypes.data_as(capi._D)))
        return out

    def tune(self, key: str, value: int) -> None:
        capi._check(capi.lib().hsv_batch_tune(self._h, key.encode(), int(value)))

    def info(self) -> dict:
        ms = ctypes.c_double()
        st = ctypes.c_int()
        amps = ctypes.c_uint64()
        capi._check(capi.lib().hsv_batch_info(self._h, ctypes.byref(ms), ctypes.byref(st), ctypes.byref(amps)))
        return {"last_ms": ms.value, "stages": st.value, "pass_amplitudes": amps.value}


def ansatz(n: int, layers: int, angles: np.ndarray) -> List[Gate]:
    """Hardware-efficient ansatz: per layer RY then RZ on every qubit, then CZ
    along a chain; ``angles`` has 2 * n * layers entries."""
    g: List[Gate] = []
    a = iter(angles)
    for _ in range(layers):
        g += [Gate("RY", (q,), (), (float(next(a)),)) for q in range(n)]
        g += [Gate("RZ", (q,), (), (float(next(a)),)) for q in range(n)]
        g += [Gate("Z", (q + 1,), (q,)) for q in range(n - 1)]
    return g


def make_batch(num: int = 1024, seed: int = 0, n_range=(8, 20), layers=(2, 6)):
    """Parameter sets of the ansatz: instance i has n ~ U{n_range}, layers ~
    U{layers} and angles ~ U[0, 2 pi)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(num):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        L = int(rng.integers(layers[0], layers[1] + 1))
        out.append((n, ansatz(n, L, rng.uniform(0, 2 * math.pi, 2 * n * L))))
    return out


CANDIDATES = [{"chunk_bits": c, "threads": t, "grid_per_cu": g}
              for c in (10, 11, 12) for t in (128, 256, 512) for g in (4, 8, 16)]


def autotune(batch: CircuitBatch, candidates=None, reps: int = 2) -> dict:
    """Times every shape on this batch and leaves the fastest set."""
    best, best_ms, table = None, float("inf"), []
    for cand in candidates or CANDIDATES:
        try:
            for k, v in cand.items():
                batch.tune(k, v)
            batch.run()
            ms = min(_timed(batch) for _ in range(reps))
        except capi.HsvError:
            continue
        table.append(dict(cand, ms=ms))
        if ms < best_ms:
            best, best_ms = cand, ms
    for k, v in best.items():
        batch.tune(k, v)
    return {"best": best, "best_ms": best_ms, "table": table}


def _timed(batch: CircuitBatch) -> float:
    batch.run()
    return batch.info()["last_ms"]


def apply_tuning(batch: CircuitBatch) -> None:
    if os.path.exists(TUNING_FILE):
        with open(TUNING_FILE) as f:
            best = json.load(f).get("best") or {}
        for k, v in best.items():
            try:
                batch.tune(k, v)
            except capi.HsvError:
                pass


def _saved_shape():
    if os.path.exists(TUNING_FILE):
        with open(TUNING_FILE) as f:
            return json.load(f).get("best")
    return None


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(description="batched small circuits")
    ap.add_argument("--instances", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--dtype", default="complex128", choices=["complex128", "complex64"])
    ap.add_argument("--tune", action="store_true", help="autotune the shape first")
    ap.add_argument("--tuning-out", default="")
    args = ap.parse_args(argv)
    insts = make_batch(args.instances, args.seed)
    S = 16 if args.dtype == "complex128" else 8
    with CircuitBatch(insts, args.dtype) as b:
        tuning = None
        if args.tune:
            tuning = autotune(b)
            if args.tuning_out:
                with open(args.tuning_out, "w") as f:
                    json.dump(tuning, f, indent=1)
            print(f"[batch] tuned: {tuning['best']} {tuning['best_ms']:.3f} ms", file=sys.stderr, flush=True)
        else:
            apply_tuning(b)
        b.run()
        ms = [_timed(b) for _ in range(args.reps)]
        info = b.info()
        mean = float(np.mean(ms))
        nbytes = 2 * S * info["pass_amplitudes"]
        gbs = nbytes / (mean * 1e-3) / 1e9
        upd = sum(len(g) * 2 ** n for n, g in insts)
        out = {"metric": "batched small circuits", "instances": len(insts), "ms_per_batch": mean,
               "min_ms": float(np.min(ms)), "instances_per_s": len(insts) / (mean * 1e-3),
               "amplitudes": int(sum(2 ** n for n, _ in insts)), "gates": int(sum(len(g) for _, g in insts)),
               "stages": info["stages"], "gate_amp_updates_per_s": upd / (mean * 1e-3),
               "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s", "frac": gbs / 8000.0,
                            "bytes_per_batch": nbytes,
                            "per_batch": f"2 x {S} B x the amplitudes each stage moves (every segment reads and "
                                         f"writes its instance once)"},
               "tuning": (tuning or {}).get("best") or _saved_shape(), "dtype": args.dtype, "reps": args.reps}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
