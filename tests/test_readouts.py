"""Read-outs beyond the norm: probability and project against the reference
build's outputs (golden vectors), Pauli expectation values against explicit
operators and closed forms. CPU tests pin the oracle side; ``gpu`` tests run
the product through the C ABI."""
import itertools
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, rand_state
from oracle.oracle import pauli_expectation, project_kept

RD = np.load(os.path.join(GOLDEN, "ref_readouts.npz"))
NCASES = len([k for k in RD.files if k.endswith("_qubits")])

PAULI = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
         "Z": np.diag([1.0, -1.0])}


def dense_pauli(labels):
    """labels[q] for qubit q (qubit 0 = least significant index bit)."""
    m = np.array([[1.0]])
    for lab in labels:   # kron(P_q, M) puts qubit q above the ones already in M
        m = np.kron(PAULI[lab], m)
    return m


def masks(labels):
    xm = sum(1 << q for q, lab in enumerate(labels) if lab in "XY")
    zm = sum(1 << q for q, lab in enumerate(labels) if lab in "ZY")
    return xm, zm


def ghz(n):
    s = np.zeros(2 ** n, complex)
    s[0] = s[-1] = 1 / math.sqrt(2)
    return s


# ---- CPU: the oracle side pinned ----------------------------------------------------
@pytest.mark.parametrize("i", range(NCASES))
def test_project_selection_matches_reference(i):
    """project_kept / p equals the reference's projected state (which drops
    the projected qubits and divides by p, state_vector.hpp:829-895)."""
    a, qs, vs, want = (RD[f"proj{i}_{k}"] for k in ("in", "qubits", "values", "out"))
    kept = project_kept(a, qs, vs)
    p = float(np.sum(np.abs(kept) ** 2))
    assert np.abs(kept / p - want).max() <= 1e-14


def test_probability_golden_is_the_marginal():
    a, prob = RD["prob_in"], RD["prob"]
    idx = np.arange(a.size)
    for q in range(int(RD["n"])):
        p1 = float(np.sum(np.abs(a[((idx >> q) & 1) == 1]) ** 2))
        assert abs(prob[q][1] - p1) <= 1e-14 and abs(prob[q][0] - (1 - p1)) <= 1e-14


def test_pauli_expectation_matches_dense_operators():
    n = 4
    a = rand_state(n, 3)
    for labels in itertools.product("IXYZ", repeat=n):
        want = np.vdot(a, dense_pauli(labels) @ a)
        got = pauli_expectation(a, *masks(labels))
        assert abs(got - want) <= 1e-13


def test_pauli_expectation_closed_forms():
    n = 6
    s = ghz(n)
    assert abs(pauli_expectation(s, 0, 0b11) - 1) <= 1e-15          # Z0 Z1
    assert abs(pauli_expectation(s, 0, 0b1)) <= 1e-15               # Z0
    assert abs(pauli_expectation(s, (1 << n) - 1, 0) - 1) <= 1e-15  # X^n
    # Y0 Y1 X2..X5 = i^2 * (X-string with Z on 0,1): on GHZ, <Y Y X X X X> = -1
    assert abs(pauli_expectation(s, (1 << n) - 1, 0b11) + 1) <= 1e-15


def test_pauli_masks_parser():
    from hisvsim_amd.capi import pauli_masks
    assert pauli_masks("X0 Z3 Y5") == ((1 << 0) | (1 << 5), (1 << 3) | (1 << 5))
    assert pauli_masks("") == (0, 0) and pauli_masks("I4") == (0, 0)
    for bad in ("Q1", "X1 Z1"):
        with pytest.raises(ValueError):
            pauli_masks(bad)


# ---- GPU: the product --------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("i", range(NCASES))
def test_gpu_project_matches_reference(i, dtype):
    """Product keeps all qubits and renormalises by 1/sqrt(p): its kept
    amplitudes equal the reference's projected state times sqrt(p)."""
    from hisvsim_amd import capi
    a, qs, vs, want = (RD[f"proj{i}_{k}"] for k in ("in", "qubits", "values", "out"))
    n = int(RD["n"])
    tol = 1e-14 if dtype == "complex128" else 2e-7
    with capi.State(n, dtype) as s:
        s.write(a)
        p = s.project(list(qs), list(vs))
        got = s.read()
        assert abs(s.norm2() - 1) <= 10 * tol
    kept_p = float(np.sum(np.abs(project_kept(a, qs, vs)) ** 2))
    assert abs(p - kept_p) <= tol
    idx = np.arange(2 ** n)
    keep = np.ones(2 ** n, bool)
    for q, v in zip(qs, vs):
        keep &= ((idx >> q) & 1) == v
    assert np.abs(got[~keep]).max() == 0.0
    assert np.abs(got[keep] - want * math.sqrt(p)).max() <= 10 * tol


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_gpu_renormalize(dtype):
    from hisvsim_amd import capi
    n = 10
    a = rand_state(n, 21)
    tol = 1e-15 if dtype == "complex128" else 1e-7
    with capi.State(n, dtype) as s:
        s.write(0.5 * a)
        assert abs(s.renormalize() - 0.25) <= 4 * tol
        assert np.abs(s.read() - a).max() <= 4 * tol
        assert abs(s.norm2() - 1) <= 10 * tol
        s.write(np.zeros(2 ** n))
        with pytest.raises(capi.HsvError):
            s.renormalize()


@pytest.mark.gpu
def test_gpu_project_zero_probability_is_an_error():
    from hisvsim_amd import capi
    with capi.State(5) as s:          # |00000>: P(qubit 2 = 1) = 0
        before = s.read()
        with pytest.raises(capi.HsvError):
            s.project([2], [1])
        assert np.array_equal(s.read(), before)
        with pytest.raises(ValueError):
            s.project([1, 1], [0, 0])


@pytest.mark.gpu
def test_gpu_probability_matches_reference():
    from hisvsim_amd import capi
    a, prob = RD["prob_in"], RD["prob"]
    with capi.State(int(RD["n"])) as s:
        s.write(a)
        for q in range(int(RD["n"])):
            for v in (0, 1):
                assert abs(s.probability(q, v) - prob[q][v]) <= 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_gpu_pauli_expectation(dtype):
    from hisvsim_amd import capi
    n = 12
    a = rand_state(n, 9)
    rng = np.random.default_rng(4)
    tol = 1e-13 if dtype == "complex128" else 1e-6
    with capi.State(n, dtype) as s:
        s.write(a)
        ref = s.read()    # the state as stored (rounded to complex64 if so)
        for _ in range(40):
            labels = "".join(rng.choice(list("IXYZ"), size=n))
            xm, zm = masks(labels)
            want = pauli_expectation(ref, xm, zm)
            assert abs(want.imag) <= 1e-12
            assert abs(s.expectation((xm, zm)) - want.real) <= tol
        assert abs(s.expectation("") - s.norm2()) <= tol
        with pytest.raises(capi.HsvError):
            s.expectation((1 << n, 0))


@pytest.mark.gpu
def test_gpu_ghz_expectations():
    """GHZ built on the device (H, CX chain): <Z_i Z_j> = 1, <Z_i> = 0,
    <X^n> = 1."""
    from hisvsim_amd import capi
    from hisvsim_amd.circuit import Gate
    n = 20
    with capi.State(n) as s:
        s.apply_gates([Gate("H", (0,))] + [Gate("X", (q,), (q - 1,)) for q in range(1, n)])
        assert abs(s.expectation("Z0 Z19") - 1) <= 1e-12
        assert abs(s.expectation("Z7")) <= 1e-12
        assert abs(s.expectation(" ".join(f"X{q}" for q in range(n))) - 1) <= 1e-12
        assert abs(s.project([5], [1]) - 0.5) <= 1e-12
        want = np.zeros(2 ** n, complex)
        want[-1] = 1
        assert np.abs(s.read() - want).max() <= 1e-12
