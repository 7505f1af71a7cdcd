"""Per-run accuracy checks of a simulated state.

Counterpart of the reference driver's ``verify_compute``
(qasm_assembler_standalone.py:716-1198), which runs a circuit and compares
the result with the initial state (``sim == sim_init``, meant for circuits
followed by their inverse). Here a run can be checked two ways:

- against an FP64 reference: the same circuit on the FP64 per-gate path in
  a second state, reporting the infidelity 1 - |<a|b>|^2 / (|a|^2 |b|^2)
  and max |a - b|;
- by its inverse: the inverse circuit (``inverse_circuit``) applied to the
  result on the FP64 per-gate path must lead back to |0...0>; reported as
  the infidelity with |0...0> and the 2-norm distance to it.

Inverses are exact: a gate whose parameters negated (or the gate itself)
give its inverse matrix is inverted that way, the reference's RXX by three
RXX gates, any other one-target gate becomes the general C1 gate with the
conjugate transpose of its matrix (the same controls). ``C2`` has no
inverse here: the reference's C2 constructor repeats column 0
(basic_gates.hpp:972-987), so its matrix is in general singular.
"""
from __future__ import annotations

from typing import List, Sequence

import numpy as np

from . import capi
from .circuit import Gate

_TOL = 1e-13
FRAC = {"RXFRAC", "RYFRAC", "RZFRAC", "RIFRAC", "R1FRAC"}


def _matrix(g: Gate) -> np.ndarray:
    return capi.gate_matrix(Gate(g.name, g.targets, (), g.params))


def _seq_matrix(seq: Sequence[Gate]) -> np.ndarray:
    m = None
    for g in seq:
        x = _matrix(g)
        m = x if m is None else x @ m
    return m


def _is_inverse(m: np.ndarray, seq: Sequence[Gate]) -> bool:
    try:
        c = _seq_matrix(seq)
    except (ValueError, capi.HsvError):
        return False
    return c.shape == m.shape and np.abs(c @ m - np.eye(m.shape[0])).max() <= _TOL


def _negated(g: Gate, params: tuple) -> Gate:
    return Gate(g.name, g.targets, g.controls, params)


def inverse_gate(g: Gate) -> List[Gate]:
    """Gates (one, or three for RXX) whose product is the inverse of g's
    matrix, on g's qubits with g's controls, in the order to apply them."""
    name = g.name.upper()
    m = _matrix(g)
    cands = [[g]]
    if name in FRAC:
        cands.append([_negated(g, (-g.params[0], g.params[1]))])
    elif g.params:
        neg = tuple(-complex(p) if isinstance(p, complex) else -p for p in g.params)
        cands.append([_negated(g, neg)])
        if name == "RXX":
            # the reference's RXX(t) is exp(-i t/2 X(x)X) times a fixed
            # involution Q = RXX(0) (basic_gates.hpp:591), so its inverse is
            # Q RXX(-t) Q
            q = _negated(g, (0.0,))
            cands.append([q, _negated(g, neg), q])
    if name == "U3":
        th, ph, lam = g.params
        cands.append([Gate("U3", g.targets, g.controls, (-th, -lam, -ph))])
    if name == "U2":
        ph, lam = g.params
        cands.append([Gate("U3", g.targets, g.controls, (-np.pi / 2, -lam, -ph))])
    if len(g.targets) == 1:
        u = m.conj().T
        cands.append([Gate("C1", g.targets, g.controls, tuple(complex(x) for x in u.reshape(-1)))])
    for c in cands:
        if _is_inverse(m, c):
            return c
    raise ValueError(f"no exact inverse for {g.name}")


def inverse_circuit(gates: Sequence[Gate]) -> List[Gate]:
    """The inverse circuit: each gate's inverse, in reverse order (identity
    gates dropped)."""
    out: List[Gate] = []
    for g in reversed(gates):
        if g.name.upper() != "I":
            out.extend(inverse_gate(g))
    return out


def against_fp64(s: "capi.State", gates: Sequence[Gate]) -> dict:
    """The state s against the circuit run from |0...0> on the FP64 per-gate
    path in a second state of s's size and storage type."""
    try:
        ref = capi.State(s.n, s.dtype)
    except capi.HsvError as e:
        return {"check": "fp64", "skipped": f"no room for a second state ({e})"}
    with ref:
        ref.apply_gates(gates)
        ref.synchronize()
        ov = ref.inner(s)
        na, nb = s.norm2(), ref.norm2()
        return {"check": "fp64", "reference": "FP64 per-gate path, same circuit, second state",
                "infidelity": max(0.0, 1.0 - abs(ov) ** 2 / (na * nb)), "max_abs_err": ref.max_abs_diff(s),
                "norm2": na}


def by_inverse(s: "capi.State", gates: Sequence[Gate], tol: float = 1e-6) -> dict:
    """Applies the inverse circuit to s (FP64 per-gate path, s is consumed)
    and compares the result with |0...0>: infidelity 1 - |a_0|^2 / |a|^2,
    2-norm distance sqrt(|a|^2 - 2 Re a_0 + 1), verified when the
    infidelity is at most tol."""
    s.apply_gates(inverse_circuit(gates))
    s.synchronize()
    a0 = complex(s.read(0, 1)[0])
    n2 = s.norm2()
    infid = max(0.0, 1.0 - abs(a0) ** 2 / n2)
    return {"check": "inverse", "reference": "inverse circuit on the FP64 per-gate path, back to |0...0>",
            "infidelity": infid, "l2_err": float(np.sqrt(max(0.0, n2 - 2 * a0.real + 1.0))),
            "verified": infid <= tol}
