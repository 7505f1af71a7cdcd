"""Gate records and the benchmark circuit generators.

A gate is ``Gate(name, targets, controls, params)``: ``name`` is a reference
gate name (basic_gates.hpp:1487-1522, case-insensitive), ``targets`` and
``controls`` are qubit indices (qubit q = bit q of the amplitude index), and
``params`` are the complex constructor arguments of the reference aggregate
(``AggregateGate_1..3``, ``_C1``, ``_C2``) or ``(numerator, power)`` for the
``*Frac`` gates. A controlled gate is the reference's BoundGate with controls
(state_vector.hpp:1339-1347), e.g. CX = ``Gate("X", (t,), (c,))``.

The generators build the workloads the simulator is measured on; each is
seeded with ``numpy.random.default_rng(seed)``.
"""
from __future__ import annotations

import math
from typing import List, NamedTuple, Sequence

import numpy as np


class Gate(NamedTuple):
    name: str
    targets: tuple
    controls: tuple = ()
    params: tuple = ()

    def qubits(self) -> tuple:
        return tuple(self.controls) + tuple(self.targets)


TWO_TARGET = {"SWAP", "C2", "RXX", "RYY", "RZZ"}
DIAGONAL = {"Z", "S", "T", "SDG", "TDG", "RZ", "U1", "RI", "R1", "RZZ", "I",
            "RZFRAC", "RIFRAC", "R1FRAC"}


def bernstein_vazirani(n: int, seed: int = 0):
    """BV pattern: data qubits 0..n-2, ancilla n-1, seed-drawn hidden mask.

    X, H on the ancilla; H on every data qubit; CX(data_i -> ancilla) for each
    set bit of the mask; H on the data qubits. From |0>, the result is
    amp[mask] = 1/sqrt2 and amp[mask + 2^(n-1)] = -1/sqrt2.
    Returns (gates, mask).
    """
    rng = np.random.default_rng(seed)
    nd = n - 1
    mask = int(rng.integers(0, 2 ** nd))
    anc = n - 1
    g: List[Gate] = [Gate("X", (anc,)), Gate("H", (anc,))]
    g += [Gate("H", (q,)) for q in range(nd)]
    g += [Gate("X", (anc,), (q,)) for q in range(nd) if (mask >> q) & 1]
    g += [Gate("H", (q,)) for q in range(nd)]
    return g, mask


def qft(n: int) -> List[Gate]:
    """QFT structure: for each qubit j, H(j) then controlled-phase(pi/2^(m-j))
    from every later qubit m (the reference's R1 = diag(1, e^{i theta}),
    basic_gates.hpp:716-728, controlled), then the qubit-reversal swaps."""
    g: List[Gate] = []
    for j in range(n):
        g.append(Gate("H", (j,)))
        for m in range(j + 1, n):
            g.append(Gate("R1", (j,), (m,), (math.pi / 2 ** (m - j),)))
    for j in range(n // 2):
        g.append(Gate("SWAP", (j, n - 1 - j)))
    return g


def gaussian_state(n: int, seed: int = 0, dtype=np.complex128) -> np.ndarray:
    """i.i.d. complex Gaussian amplitudes, normalised."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    a /= np.linalg.norm(a)
    return a.astype(dtype)


def random_layered(n: int, depth: int, seed: int = 0) -> List[Gate]:
    """Each layer: U3(theta, phi, lambda) on every qubit with angles uniform in
    [0, 2pi), then CZ on a random perfect matching (``rng.permutation(n)``
    paired consecutively; with odd n the last qubit idles)."""
    rng = np.random.default_rng(seed)
    g: List[Gate] = []
    for _ in range(depth):
        ang = rng.uniform(0.0, 2 * math.pi, size=(n, 3))
        for q in range(n):
            g.append(Gate("U3", (q,), (), tuple(float(x) for x in ang[q])))
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            a, b = int(perm[i]), int(perm[i + 1])
            g.append(Gate("Z", (max(a, b),), (min(a, b),)))
    return g


def random_layered_layers(n: int, depth: int, seed: int = 0):
    """Same circuit as :func:`random_layered`, returned per layer as
    (u3_gates, cz_pairs) so partitioners can use the matching structure."""
    gates = random_layered(n, depth, seed)
    per = n + n // 2
    out = []
    for L in range(depth):
        chunk = gates[L * per:(L + 1) * per]
        out.append((chunk[:n], [(g.controls[0], g.targets[0]) for g in chunk[n:]]))
    return out


def gate_amplitudes(g: Gate, n: int) -> int:
    """Amplitudes a per-gate pass must rewrite (the algorithmic work unit):
    2^(n-c) for a dense gate with c controls; half of that for a phase gate on
    the |1> half (Z, S, T, SDG, TDG); SWAP rewrites half."""
    c = len(g.controls)
    name = g.name.upper()
    if name in ("Z", "S", "T", "SDG", "TDG"):
        return 2 ** (n - c - 1)
    if name == "SWAP":
        return 2 ** (n - c - 1)
    if name == "I":
        return 0
    return 2 ** (n - c)


def part_qubits(gates: Sequence[Gate]) -> tuple:
    qs = set()
    for g in gates:
        qs.update(g.qubits())
    return tuple(sorted(qs))
