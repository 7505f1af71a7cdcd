"""Batched QAOA MaxCut on the GPU (libhsv's hsv_qaoa_* entry points).

Instances: n ~ U{12..22} qubits, a random 3-regular graph (an odd n gets one
vertex of degree 2), p ~ U{1..4} layers, gamma and beta ~ U[0, pi) per
layer. Circuit: H on every qubit, then per layer ZZ(gamma) per edge and
RX(2 beta) per qubit, where ZZ(gamma) = exp(-i gamma/2 Z x Z) is the
reference's RZZ (basic_gates.hpp:1249-1252). Output: <C>, with
C = sum over edges of (1 - Z_u Z_v) / 2 (the cut size).
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import time
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from . import capi
from .circuit import Gate


@dataclass
class Instance:
    n: int
    edges: List[tuple]
    gammas: List[float]
    betas: List[float]

    @property
    def p(self) -> int:
        return len(self.gammas)


def random_3regular(n: int, rng: np.random.Generator) -> List[tuple]:
    """Configuration model with rejection: 3 stubs per vertex (one vertex
    gets 2 when n is odd), paired at random until the graph is simple."""
    deg = [3] * n
    if n % 2:
        deg[int(rng.integers(n))] = 2
    stubs = np.repeat(np.arange(n), deg)
    for _ in range(10000):
        perm = rng.permutation(stubs)
        pairs = perm.reshape(-1, 2)
        if np.any(pairs[:, 0] == pairs[:, 1]):
            continue
        es = {tuple(sorted((int(a), int(b)))) for a, b in pairs}
        if len(es) == len(pairs):
            return sorted(es)
    raise RuntimeError("could not draw a simple 3-regular graph")


def make_batch(num: int = 512, seed: int = 0, n_range=(12, 22), p_range=(1, 4)) -> List[Instance]:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(num):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        edges = random_3regular(n, rng)
        p = int(rng.integers(p_range[0], p_range[1] + 1))
        g = [float(x) for x in rng.uniform(0, math.pi, p)]
        b = [float(x) for x in rng.uniform(0, math.pi, p)]
        out.append(Instance(n, edges, g, b))
    return out


def circuit(inst: Instance) -> List[Gate]:
    """The instance as reference gates (for the CPU oracle)."""
    gs = [Gate("H", (q,)) for q in range(inst.n)]
    for g, b in zip(inst.gammas, inst.betas):
        gs += [Gate("RZZ", (u, v), (), (g,)) for u, v in inst.edges]
        gs += [Gate("RX", (q,), (), (2 * b,)) for q in range(inst.n)]
    return gs


def cut_values(n: int, edges: Sequence[tuple]) -> np.ndarray:
    x = np.arange(2 ** n, dtype=np.int64)
    c = np.zeros(2 ** n, dtype=np.int64)
    for u, v in edges:
        c += ((x >> u) ^ (x >> v)) & 1
    return c


def expectation_p1(n: int, edges: Sequence[tuple], gamma: float, beta: float) -> float:
    """Closed form of <C> for one QAOA layer on MaxCut (Wang, Hadfield, Jiang,
    Rieffel, Phys. Rev. A 97, 022304 (2018), Theorem 1), written for
    U(C, g) = exp(-i g C). Here ZZ(gamma) = exp(-i gamma/2 Z Z) =
    e^{-i gamma/2} exp(+i gamma C_uv), so g = -gamma."""
    g = -gamma
    adj = [set() for _ in range(n)]
    for u, v in edges:
        adj[u].add(v)
        adj[v].add(u)
    total = 0.0
    for u, v in edges:
        d, e = len(adj[u]) - 1, len(adj[v]) - 1
        f = len(adj[u] & adj[v])
        term = 0.5 + 0.25 * math.sin(4 * beta) * math.sin(g) * (math.cos(g) ** d + math.cos(g) ** e)
        term -= 0.25 * math.sin(2 * beta) ** 2 * math.cos(g) ** (d + e - 2 * f) * (1 - math.cos(2 * g) ** f)
        total += term
    return total


class QAOABatch:
    """A batch of instances on the GPU, run as one hipGraph."""

    def __init__(self, instances: Sequence[Instance]):
        self.instances = list(instances)
        L = capi.lib()
        num = len(self.instances)
        nq = (ctypes.c_int * num)(*[i.n for i in self.instances])
        offs = [0]
        flat = []
        for i in self.instances:
            offs.append(offs[-1] + len(i.edges))
            for u, v in i.edges:
                flat += [u, v]
        eo = (ctypes.c_int * (num + 1))(*offs)
        ed = (ctypes.c_int * max(1, len(flat)))(*flat)
        ly = (ctypes.c_int * num)(*[i.p for i in self.instances])
        self._h = ctypes.c_void_p()
        capi._check(L.hsv_qaoa_create(num, nq, eo, ed, ly, ctypes.byref(self._h)))
        self.set_angles([i.gammas for i in self.instances], [i.betas for i in self.instances])

    def close(self):
        if self._h:
            capi.lib().hsv_qaoa_destroy(self._h)
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

    def set_angles(self, gammas: Sequence[Sequence[float]], betas: Sequence[Sequence[float]]) -> None:
        g = np.ascontiguousarray(np.concatenate([np.asarray(x, float) for x in gammas]) if gammas else np.zeros(1))
        b = np.ascontiguousarray(np.concatenate([np.asarray(x, float) for x in betas]) if betas else np.zeros(1))
        capi._check(capi.lib().hsv_qaoa_set_angles(self._h, g.ctypes.data_as(capi._D), b.ctypes.data_as(capi._D)))

    def run(self) -> np.ndarray:
        out = np.zeros(len(self.instances))
        capi._check(capi.lib().hsv_qaoa_run(self._h, out.ctypes.data_as(capi._D)))
        return out

    def last_ms(self) -> float:
        v = ctypes.c_double()
        capi._check(capi.lib().hsv_qaoa_last_ms(self._h, ctypes.byref(v)))
        return v.value

    def tune(self, key: str, value: int) -> None:
        capi._check(capi.lib().hsv_qaoa_tune(self._h, key.encode(), int(value)))

    def state(self, i: int) -> np.ndarray:
        out = np.zeros(2 * 2 ** self.instances[i].n)
        capi._check(capi.lib().hsv_qaoa_read_state(self._h, i, out.ctypes.data_as(capi._D)))
        return out[0::2] + 1j * out[1::2]

    def total_amplitudes(self) -> int:
        v = ctypes.c_uint64()
        capi._check(capi.lib().hsv_qaoa_total_amplitudes(self._h, ctypes.byref(v)))
        return v.value

    def bytes_per_run(self, chunk_bits: int = 12) -> int:
        """HBM bytes of one run as the schedule moves them, 8 B per complex64
        amplitude: the initial write, per layer one read and one write of
        the state in stage A (the ZZ layer and RX on the chunk_bits low
        qubits) and, for instances wider than that, one more in stage B (RX
        on the high qubits), and the read of the cost reduction."""
        tot = 0
        for i in self.instances:
            passes = 0.5 + i.p * (1 + (i.n > chunk_bits)) + 0.5
            tot += int(passes * 2 * 8 * 2 ** i.n)
        return tot

    def gate_updates_per_run(self) -> int:
        """Gate-amplitude updates of one run: every gate (H layer, ZZ per edge,
        RX per qubit) counted over its instance's whole state."""
        return sum((i.n + i.p * (len(i.edges) + i.n)) * 2 ** i.n for i in self.instances)


CANDIDATES = [{"chunk_bits": c, "threads": t, "grid_per_cu": g}
              for c in (11, 12, 13) for t in (128, 256, 512) for g in (2, 4, 8)]

TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_qaoa.json")


def autotune(batch: QAOABatch, candidates=None, reps: int = 2) -> dict:
    """Times every tile shape on this batch and leaves the fastest set."""
    best, best_ms, table = None, float("inf"), []
    for cand in candidates or CANDIDATES:
        try:
            for k, v in cand.items():
                batch.tune(k, v)
            batch.run()  # capture + warm up
            ms = min(_timed(batch) for _ in range(reps))
        except capi.HsvError:
            continue
        table.append(dict(cand, ms=ms))
        if ms < best_ms:
            best, best_ms = cand, ms
    for k, v in best.items():
        batch.tune(k, v)
    return {"best": best, "best_ms": best_ms, "table": table}


def _timed(batch: QAOABatch) -> float:
    batch.run()
    return batch.last_ms()


def load_tuning() -> dict | None:
    if os.path.exists(TUNING_FILE):
        with open(TUNING_FILE) as f:
            return json.load(f)
    return None


def apply_tuning(batch: QAOABatch) -> None:
    t = load_tuning()
    if t and t.get("best"):
        for k, v in t["best"].items():
            try:
                batch.tune(k, v)
            except capi.HsvError:
                pass


def main(argv=None) -> int:
    """Batched QAOA benchmark (512 instances by default):

        python -m hisvsim_amd.qaoa [--instances 512] [--seed 0] [--reps 20]
            [--tune] [--tuning-out FILE]

    Prints one JSON line: mean ms per batch (one hipGraph replay, inputs
    resident on the device), instances/s, algorithmic HBM GB/s of the gate
    passes, the tile shape used, and the largest error of the p = 1
    instances against the closed form."""
    import argparse
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--tune", action="store_true", help="autotune the tile shape first")
    ap.add_argument("--tuning-out", default="", help="where to write the tuning result (with --tune)")
    args = ap.parse_args(argv)
    insts = make_batch(args.instances, args.seed)
    with QAOABatch(insts) as b:
        tuning = None
        if args.tune:
            tuning = autotune(b)
            if args.tuning_out:
                with open(args.tuning_out, "w") as f:
                    json.dump(tuning, f, indent=1)
            print(f"[qaoa] tuned: {tuning['best']} {tuning['best_ms']:.3f} ms", file=sys.stderr, flush=True)
        else:
            apply_tuning(b)
        vals = b.run()  # capture + warm-up
        ms = []
        for _ in range(args.reps):
            vals = b.run()
            ms.append(b.last_ms())
        mean_ms = float(np.mean(ms))
        err = max((abs(vals[i] - expectation_p1(x.n, x.edges, x.gammas[0], x.betas[0]))
                   for i, x in enumerate(insts) if x.p == 1), default=0.0)
        shape = tuning["best"] if tuning else (load_tuning() or {}).get("best") or {}
        nbytes = b.bytes_per_run(int(shape.get("chunk_bits", 12)))
        gbs = nbytes / (mean_ms * 1e-3) / 1e9
        out = {"metric": "batched QAOA MaxCut <C>", "instances": len(insts), "ms_per_batch": mean_ms,
               "min_ms": float(np.min(ms)), "instances_per_s": len(insts) / (mean_ms * 1e-3),
               "amplitudes": b.total_amplitudes(),
               "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s", "frac": gbs / 8000.0,
                            "bytes_per_batch": nbytes,
                            "per_batch": "state bytes x passes of the schedule (init write, stage A and B read + "
                                         "write per layer, cost read), 8 B per complex64 amplitude"},
               "gate_amp_updates_per_s": b.gate_updates_per_run() / (mean_ms * 1e-3),
               "p1_instances": sum(1 for x in insts if x.p == 1), "p1_max_abs_err": err,
               "tuning": tuning["best"] if tuning else (load_tuning() or {}).get("best"),
               "dtype": "complex64 state, FP64 reductions", "reps": args.reps}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
