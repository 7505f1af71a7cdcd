#!/usr/bin/env python3
"""Writes the golden vectors in tests/golden/ from the reference simulator.

Runs only where the reference checkout exists: it uses oracle/_ref/libref.so
(``make -C oracle ref``), the reference's own headers compiled behind a small
C harness, and calls its per-gate path (BoundGate::apply,
state_vector.hpp:1349) and its live part driver (gather_and_execute_on_multi,
execute.hpp:318-470). Inputs are seeded; outputs are the reference's
amplitudes. Nothing here is imported by the product.

    python tests/golden/make_golden.py
"""
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from gatesets import ALL_ONE, ALL_TWO, params_for, random_circuit  # noqa: E402
from hisvsim_amd.circuit import Gate, bernstein_vazirani, qft, random_layered  # noqa: E402
from hisvsim_amd.partition import greedy_parts  # noqa: E402
from oracle.oracle import RefSim  # noqa: E402


def rand_state(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    return a / np.linalg.norm(a)


def spec(g):
    return {"name": g.name, "targets": list(g.targets), "controls": list(g.controls),
            "params": [[complex(p).real, complex(p).imag] for p in g.params]}


def gates_fixture(ref):
    """Every reference gate type at several target/control placements, n=5."""
    n = 5
    rng = np.random.default_rng(2024)
    specs, ins, outs = [], [], []
    placements1 = [((0,), ()), ((3,), ()), ((1,), (4,)), ((4,), (0, 2))]
    placements2 = [((0, 1), ()), ((1, 0), ()), ((3, 1), ()), ((2, 4), (0,))]
    for name in ALL_ONE + ALL_TWO:
        for t, c in (placements2 if name in ALL_TWO else placements1):
            g = Gate(name, t, c, params_for(name, rng))
            a = rand_state(n, len(specs))
            b = a.copy()
            ref.apply_gates(b, [g])
            s = spec(g)
            if name.endswith("FRAC"):
                s["params"] = [int(g.params[0]), int(g.params[1])]
            specs.append(s)
            ins.append(a)
            outs.append(b)
    return {"specs": json.dumps(specs), "n": n, "inputs": np.array(ins), "outputs": np.array(outs)}


def circuits_fixture(ref):
    out = {}
    # Bernstein-Vazirani pattern, n=12, from |0>
    n = 12
    gates, mask = bernstein_vazirani(n, seed=3)
    a = np.zeros(2 ** n, complex); a[0] = 1
    ref.apply_gates(a, gates)
    out["bv12_mask"] = mask
    out["bv12"] = a
    # QFT structure (controlled R1), n=10, Gaussian input
    n = 10
    x = rand_state(n, 7)
    y = x.copy()
    ref.apply_gates(y, qft(n))
    out["qft10_in"] = x
    out["qft10"] = y
    # QFT as the reference driver lowers cu1 (controlled U1 = controlled RZ)
    g_ref = [Gate("RZ" if g.name == "R1" else g.name, g.targets, g.controls, g.params) for g in qft(n)]
    y2 = x.copy()
    ref.apply_gates(y2, g_ref)
    out["qft10_cu1"] = y2
    # random layered, n=12 depth 6: per gate and through the part driver
    n = 12
    gates = random_layered(n, 6, seed=11)
    a = np.zeros(2 ** n, complex); a[0] = 1
    b = a.copy()
    ref.apply_gates(a, gates)
    parts = greedy_parts(gates, 6)
    ref.run_parts(b, parts, opt_slots=1)
    out["rl12"] = a
    out["rl12_parts"] = b
    out["rl12_part_sizes"] = np.array([len(p) for p in parts])
    # every gate type with up to 2 controls, n=10, per gate and in parts of
    # consecutive gates (no I: the reference's part executor aborts on it)
    n = 10
    rng = np.random.default_rng(5)
    names = [g for g in ALL_ONE + ALL_TWO if g != "I"]
    gates = random_circuit(rng, n, 120, names=names, max_controls=2)
    x = rand_state(n, 9)
    a, b = x.copy(), x.copy()
    ref.apply_gates(a, gates)
    parts = [gates[i:i + 10] for i in range(0, len(gates), 10)]
    ref.run_parts(b, parts, opt_slots=1)
    out["mix10_in"] = x
    out["mix10_specs"] = json.dumps([spec(g) if not g.name.endswith("FRAC") else
                                     dict(spec(g), params=[int(g.params[0]), int(g.params[1])]) for g in gates])
    out["mix10"] = a
    out["mix10_parts"] = b
    # n=20 random layered depth 3: checksums and sampled amplitudes
    n = 20
    gates = random_layered(n, 3, seed=21)
    a = np.zeros(2 ** n, complex); a[0] = 1
    ref.apply_gates(a, gates)
    idx = np.random.default_rng(0).choice(2 ** n, size=64, replace=False)
    out["rl20_idx"] = idx
    out["rl20_amps"] = a[idx]
    out["rl20_sum"] = a.sum()
    out["rl20_wsum"] = np.sum(np.abs(a) ** 2 * np.arange(2 ** n))
    return out


PROJECT_CASES = [((0,), (0,)), ((3,), (1,)), ((7,), (1,)), ((1, 5), (1, 0)), ((6, 2, 4), (0, 1, 1)), ((2,), ())]


def readouts_fixture(ref):
    """StateVector::probability (every qubit, both values) and ::project
    (single and several qubits; an omitted value defaults to 0) on seeded
    random states, n=8."""
    n = 8
    out = {"n": n}
    for i, (qs, vs) in enumerate(PROJECT_CASES):
        a = rand_state(n, 500 + i)
        out[f"proj{i}_in"] = a
        out[f"proj{i}_qubits"] = np.array(qs)
        out[f"proj{i}_values"] = np.array(list(vs) + [0] * (len(qs) - len(vs)))
        out[f"proj{i}_out"] = ref.project(a, qs, list(vs) + [0] * (len(qs) - len(vs)))
    # probability accumulates into one shared double inside the OpenMP loop
    # (state_vector.hpp:908-913), a data race with several threads: take it
    # with one thread, where it is the plain marginal
    ref.set_threads(1)
    a = rand_state(n, 600)
    out["prob_in"] = a
    out["prob"] = np.array([[ref.probability(a, q, v) for v in (0, 1)] for q in range(n)])
    ref.set_threads(8)
    return out


FIXTURES = {"gates": ("ref_gates.npz", gates_fixture), "circuits": ("ref_circuits.npz", circuits_fixture),
            "readouts": ("ref_readouts.npz", readouts_fixture)}


def main():
    """``python make_golden.py [gates] [circuits] [readouts]`` (default: all)."""
    ref = RefSim()
    ref.set_threads(8)
    for key in (sys.argv[1:] or list(FIXTURES)):
        fname, fn = FIXTURES[key]
        np.savez_compressed(os.path.join(HERE, fname), **fn(ref))
        print(fname, os.path.getsize(os.path.join(HERE, fname)), "bytes")


if __name__ == "__main__":
    main()
