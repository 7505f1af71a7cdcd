"""Learned cost model: a small MLP that predicts a part's kernel time from its
features, trained on the GPU (libhsv's hsv_mlp_* kernels) on per-part timings
the simulator records itself. The planner uses it to choose the part width k
and the partition; the plain partitioner (natural order, the reference's
``*_part_nat`` method) is the fallback, and measured runtime judges the
model.

Features of a part, for a state of n qubits and a kernel mode (register
part kernel, fused FP64 unitary, fused BF16-split unitary - one block of
5..7 qubits, or a 12/13-qubit tile of 5-qubit steps for wider parts - or a
part of the moving-layout chain, whose tile adds the part's low-slot
qubits):
  part qubits k, the width the device runs it at (k plus the low qubits it
  adds, the fused path's padded width, or the tile width), gate count,
  1-qubit and multi-qubit gate counts, the multi-qubit fraction, min / max /
  spread of the qubit indices, whether it touches the top 3 (high-order)
  qubits, two mode flags, n, and the number of fused blocks the pass
  applies (0 for the register kernel, 1 for a 5..7-qubit unitary, the
  5-qubit steps of a wide part).
Target: log(ms per pass scaled to a 2^30-amplitude state), so one model
covers every n (a pass is memory-bound work proportional to 2^n). A chain
plan adds its layout restore, priced per 2^30 amplitudes by the median of
the restores timed during collection.
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import time
from typing import Callable, Dict, List, Sequence, Tuple

import numpy as np

from . import capi
from .circuit import Gate
from .partition import chain_parts, closure_parts, dfs_parts, greedy_parts, nat_parts

MODES = ("part", "fused64", "fusedbf16", "chain")
FUSED_PRECISION = {"fused64": "f64", "fusedbf16": "bf16split"}
LOW_SLOTS = 5
N_FEATURES = 14
RESTORE_MS_PER_2_30 = 26.0   # used when a parameter file has no measured restore cost
PARAMS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "costmodel_params.json")


def part_features(part: Sequence[Gate], n: int, mode: str, low: Sequence[int] = ()) -> np.ndarray:
    """``low``: the chain part's low-slot qubits (mode "chain" only)."""
    gq = {q for g in part for q in g.qubits()}
    qs = sorted(gq | set(low))
    k = len(qs)
    low5 = set(range(min(LOW_SLOTS, n)))
    if mode == "part":
        width, blocks = len(set(qs) | low5), 0
    elif mode == "chain" or (mode == "fusedbf16" and k > 7):
        # the wide-part kernel's tile: 12 or 13 qubits holding slots 0-4
        width = 12 if len(set(qs) | (low5 if mode != "chain" else set())) <= 12 else 13
        blocks = len(capi.part_steps(part)) if part else 0   # the kernel's fused steps
    else:
        width, blocks = max(4 if mode == "fused64" else 5, k), 1
    n1 = sum(1 for g in part if len(g.qubits()) == 1)
    n2 = len(part) - n1
    return np.array([
        k / 10.0, width / 10.0, len(part) / 20.0, n1 / 20.0, n2 / 20.0, n2 / max(1, len(part)),
        qs[0] / 30.0, qs[-1] / 30.0, (qs[-1] - qs[0]) / 30.0, float(qs[-1] >= n - 3),
        float(mode == "fused64"), float(mode in ("fusedbf16", "chain")), (n - 25) / 5.0, blocks / 4.0,
    ], dtype=np.float32)


def target_of(ms: float, n: int) -> float:
    return math.log(ms / 2.0 ** (n - 30))


def ms_of(target: float, n: int) -> float:
    return math.exp(target) * 2.0 ** (n - 30)


class MLP:
    """The cost-model network on the GPU."""

    def __init__(self, in_features: int = N_FEATURES, hidden: int = 64, seed: int = 1):
        self.in_features, self.hidden = in_features, hidden
        self._h = ctypes.c_void_p()
        capi._check(capi.lib().hsv_mlp_create(in_features, hidden, seed, ctypes.byref(self._h)))
        c = ctypes.c_uint64()
        capi._check(capi.lib().hsv_mlp_num_params(self._h, ctypes.byref(c)))
        self.num_params = c.value

    def close(self):
        if self._h:
            capi.lib().hsv_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _fp(a: np.ndarray):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def params(self) -> np.ndarray:
        out = np.zeros(self.num_params, np.float32)
        capi._check(capi.lib().hsv_mlp_get_params(self._h, self._fp(out)))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params
        capi._check(capi.lib().hsv_mlp_set_params(self._h, self._fp(p)))

    def predict(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros(len(X), np.float32)
        capi._check(capi.lib().hsv_mlp_forward(self._h, self._fp(X), len(X), self._fp(out)))
        return out

    def gradient(self, X: np.ndarray, y: np.ndarray) -> Tuple[float, np.ndarray]:
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros(self.num_params, np.float32)
        loss = ctypes.c_float()
        capi._check(capi.lib().hsv_mlp_gradient(self._h, self._fp(X), self._fp(y), len(X), ctypes.byref(loss),
                                                self._fp(g)))
        return loss.value, g

    def train(self, X: np.ndarray, y: np.ndarray, epochs: int = 200, batch: int = 512, lr: float = 3e-3,
              seed: int = 0) -> float:
        rng = np.random.default_rng(seed)
        perm = rng.permutation(len(X))
        X = np.ascontiguousarray(X[perm], np.float32)
        y = np.ascontiguousarray(y[perm], np.float32)
        loss = ctypes.c_float()
        capi._check(capi.lib().hsv_mlp_train(self._h, self._fp(X), self._fp(y), len(X), epochs, batch, lr,
                                             ctypes.byref(loss)))
        return loss.value

    def save(self, path: str = PARAMS_FILE, meta: dict | None = None) -> None:
        with open(path, "w") as f:
            json.dump({"in": self.in_features, "hidden": self.hidden, "params": self.params().tolist(),
                       "meta": meta or {}}, f)

    @classmethod
    def load(cls, path: str = PARAMS_FILE) -> "MLP":
        with open(path) as f:
            d = json.load(f)
        m = cls(d["in"], d["hidden"])
        m.set_params(np.array(d["params"], np.float32))
        m.restore_ms_per_2_30 = float(d.get("meta", {}).get("restore_ms_per_2_30", RESTORE_MS_PER_2_30))
        return m


# ---- planning ------------------------------------------------------------------
def candidate_specs() -> List[Tuple[str, int, str]]:
    """(mode, k, partitioner) candidates: the fused unitary at k = 5..7 (FP64
    or BF16-split MFMA), wide BF16-split parts at k = 8..13 and the part
    kernel at k = 8..13 (those parts already hold the 5 low qubits), with the
    DAG-aware greedy, depth-first, qubit-set closure or natural-order
    partitioner (k = 5..10 for natural order on the part kernel), and the
    moving-layout chain at k = 12, 13. k is the block width the model picks."""
    specs = [("fused64", k, name) for k in range(5, 8) for name in ("greedy", "nat")]
    specs += [("fusedbf16", k, name) for k in range(5, 8) for name in ("closure", "dfs")]
    specs += [("fusedbf16", k, "closure") for k in range(8, 14)]
    specs += [("chain", k, "chain") for k in (12, 13)]
    specs += [("part", k, name) for k in range(8, 14) for name in ("greedy", "dfs", "closure")]
    specs += [("part", k, "nat") for k in range(5, 11)]
    return specs


def build_plan(gates: Sequence[Gate], n: int, spec: Tuple[str, int, str]) -> dict:
    mode, k, name = spec
    low = range(min(LOW_SLOTS, n))
    fn = {"dfs": dfs_parts, "closure": closure_parts, "greedy": greedy_parts}
    if mode == "chain":
        parts, lows = chain_parts(gates, k, LOW_SLOTS)
        return {"mode": mode, "k": k, "partitioner": name, "parts": parts, "lows": lows}
    if name == "nat":
        parts = nat_parts(gates, k)
    elif mode == "part" or (mode == "fusedbf16" and k >= 8):
        parts = fn[name](gates, max(k, len(low)), always=low)
    else:
        parts = fn[name](gates, k)
    return {"mode": mode, "k": k, "partitioner": name, "parts": parts}


def candidate_plans(gates: Sequence[Gate], n: int) -> List[dict]:
    return [build_plan(gates, n, s) for s in candidate_specs()]


def fallback_plan(gates: Sequence[Gate], n: int, k: int = 10) -> dict:
    """The plain partitioner: natural order, k = 10, on the part kernel."""
    return {"mode": "part", "k": k, "partitioner": "nat", "parts": nat_parts(gates, k)}


def plan_features(plan: dict, n: int) -> np.ndarray:
    lows = plan.get("lows") or [()] * len(plan["parts"])
    return np.stack([part_features(p, n, plan["mode"], low) for p, low in zip(plan["parts"], lows)])


def predict_plan_ms(model: MLP, plan: dict, n: int) -> float:
    ms = float(sum(ms_of(float(t), n) for t in model.predict(plan_features(plan, n))))
    if plan["mode"] == "chain":
        ms += getattr(model, "restore_ms_per_2_30", RESTORE_MS_PER_2_30) * 2.0 ** (n - 30)
    return ms


def choose_plan(model: MLP, gates: Sequence[Gate], n: int) -> Tuple[dict, List[dict]]:
    plans = candidate_plans(gates, n)
    for p in plans:
        p["predicted_ms"] = predict_plan_ms(model, p, n)
    best = min(plans, key=lambda p: p["predicted_ms"])
    return best, plans


def default_plan(gates: Sequence[Gate], n: int) -> dict:
    """The fixed choice without the model: the moving-layout chain at k = 13
    (the benchmark's path, greedy planner)."""
    return build_plan(gates, n, ("chain", 13, "chain"))


def plan_name(plan: dict) -> str:
    return f"{plan['mode']} k={plan['k']} {plan['partitioner']}"


def compare_choice(model: "MLP", gates: Sequence[Gate], n: int, s: capi.State, top: int = 3) -> dict:
    """What the model's choice is worth on one circuit, measured: the model's
    pick and its next-ranked candidates (``top`` plans in all), the fixed
    default (chain, k = 13) and the plain partitioner (natural order, k = 10,
    part kernel) are each run on the dense state ``s`` and timed by their
    kernel times. Returns the measured and predicted ms per plan, the
    model's regret (its pick's time over the best time measured), its gains
    over the default and over the plain partitioner, the pick's plan-level
    prediction error, and the per-part prediction errors."""
    best, plans = choose_plan(model, gates, n)
    ranked = sorted(plans, key=lambda p: p["predicted_ms"])
    named = {}
    for p in ranked:
        if len(named) >= top:
            break
        named.setdefault(plan_name(p), p)
    dflt = default_plan(gates, n)
    dflt["predicted_ms"] = predict_plan_ms(model, dflt, n)
    plain = fallback_plan(gates, n)
    plain["predicted_ms"] = predict_plan_ms(model, plain, n)
    named.setdefault(plan_name(dflt), dflt)
    named.setdefault(plan_name(plain), plain)
    measured, predicted, part_err = {}, {}, []
    for plan in named.values():
        run_plan(s, plan, timed=False)   # code objects and scratch, so that no plan pays them in its timing
    for name, plan in named.items():
        times = run_plan(s, plan)
        measured[name] = float(times.sum())
        predicted[name] = float(plan["predicted_ms"])
        pred = np.array([ms_of(float(t), n) for t in model.predict(plan_features(plan, n))])
        tp = times[:len(plan["parts"])]
        part_err.extend((np.abs(pred - tp) / tp).tolist())
    pick = plan_name(best)
    fastest = min(measured.values())
    return {"pick": pick, "measured_ms": measured, "predicted_ms": predicted,
            "regret": measured[pick] / fastest,
            "gain_over_default": measured[plan_name(dflt)] / measured[pick],
            "gain_over_plain": measured[plan_name(plain)] / measured[pick],
            "plan_rel_err": abs(predicted[pick] - measured[pick]) / measured[pick],
            "rank_of_fastest": sorted(predicted, key=predicted.get).index(min(measured, key=measured.get)),
            "part_rel_err": part_err}


def run_plan(s: capi.State, plan: dict, timed: bool = True) -> np.ndarray:
    """Runs a plan on the state; returns each part's kernel time (ms), and
    for a chain plan its layout restore as one more entry."""
    if timed:
        s.set_timing(True)
        s.kernel_times()
    if plan["mode"] == "chain":
        s.apply_fused_chain(plan["parts"], plan["lows"], "bf16split")
    for p in plan["parts"] if plan["mode"] != "chain" else ():
        if plan["mode"] in FUSED_PRECISION:
            s.apply_fused(p, FUSED_PRECISION[plan["mode"]])
        else:
            s.apply_part(p, 0)
    if plan["mode"] in ("fusedbf16", "chain"):
        s.renormalize()   # not a timed launch: the per-part times stay per part
    s.synchronize()
    if not timed:
        return np.zeros(0)
    t = s.kernel_times()
    s.set_timing(False)
    return t
