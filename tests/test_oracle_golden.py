"""Pins the CPU oracle (CPU only, no GPU).

1. Golden vectors written by the reference's own headers (tests/golden/*.npz,
   made by tests/golden/make_golden.py through oracle/_ref): every gate type
   at several placements, BV, QFT (R1 and the reference's cu1 lowering), a
   random layered circuit per gate and through the part driver, a circuit of
   every gate type with controls, and sampled amplitudes at n=20.
2. The reference's disabled doctest known answers (state_vector.hpp:1396-1739).
3. Brute force: each gate as an explicit 2^n x 2^n operator built from its
   matrix written out here (textbook forms, plus the reference's documented
   deviations: U1 = RZ, the U3 global phase, RXX's extra H, C2's repeated
   columns), checked against the oracle on random states.
4. Closed forms: BV amplitudes, QFT = bit-reversed inverse DFT.
5. Invariants: part path == per-gate path bit for bit; permutation inverse.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, bit_reverse_perm, rand_state
from gatesets import ALL_ONE, ALL_TWO, params_for
from hisvsim_amd.circuit import Gate, bernstein_vazirani, qft, random_layered
from hisvsim_amd.partition import greedy_parts

I2 = np.eye(2)
HM = np.array([[1, 1], [1, -1]]) / math.sqrt(2)


def from_spec(s):
    if s["name"].upper().endswith("FRAC"):
        return Gate(s["name"], tuple(s["targets"]), tuple(s["controls"]), tuple(s["params"]))
    return Gate(s["name"], tuple(s["targets"]), tuple(s["controls"]), tuple(complex(a, b) for a, b in s["params"]))


# ---- 1. golden vectors from the reference build ---------------------------------
def test_oracle_matches_reference_per_gate(orc):
    f = np.load(os.path.join(GOLDEN, "ref_gates.npz"))
    specs = json.loads(str(f["specs"]))
    assert len(specs) == len(f["inputs"]) > 100
    for s, a, want in zip(specs, f["inputs"], f["outputs"]):
        got = a.copy()
        orc.apply_gate(got, from_spec(s))
        assert np.abs(got - want).max() <= 1e-15, s


def test_oracle_matches_reference_circuits(orc):
    f = np.load(os.path.join(GOLDEN, "ref_circuits.npz"))
    # BV
    n = 12
    gates, mask = bernstein_vazirani(n, seed=3)
    assert int(f["bv12_mask"]) == mask
    a = np.zeros(2 ** n, complex); a[0] = 1
    orc.apply_gates(a, gates)
    assert np.abs(a - f["bv12"]).max() <= 1e-15
    # QFT with R1 and with the reference driver's cu1 -> controlled RZ lowering
    x = f["qft10_in"]
    y = x.copy(); orc.apply_gates(y, qft(10))
    assert np.abs(y - f["qft10"]).max() <= 1e-14
    y = x.copy()
    orc.apply_gates(y, [Gate("RZ" if g.name == "R1" else g.name, g.targets, g.controls, g.params) for g in qft(10)])
    assert np.abs(y - f["qft10_cu1"]).max() <= 1e-14
    # random layered per gate and through the part driver
    n = 12
    gates = random_layered(n, 6, seed=11)
    a = np.zeros(2 ** n, complex); a[0] = 1
    b = a.copy()
    orc.apply_gates(a, gates)
    parts = greedy_parts(gates, 6)
    assert list(f["rl12_part_sizes"]) == [len(p) for p in parts]
    for p in parts:
        orc.run_part(b, p, 1)
    assert np.abs(a - f["rl12"]).max() <= 1e-15
    assert np.abs(b - f["rl12_parts"]).max() <= 1e-15
    # every gate type with controls, per gate and in parts
    gates = [from_spec(s) for s in json.loads(str(f["mix10_specs"]))]
    x = f["mix10_in"]
    a, b = x.copy(), x.copy()
    orc.apply_gates(a, gates)
    for i in range(0, len(gates), 10):
        orc.run_part(b, gates[i:i + 10], 1)
    assert np.abs(a - f["mix10"]).max() <= 1e-15
    assert np.abs(b - f["mix10_parts"]).max() <= 1e-15
    # n=20 sampled amplitudes and checksums
    n = 20
    a = np.zeros(2 ** n, complex); a[0] = 1
    orc.apply_gates(a, random_layered(n, 3, seed=21))
    assert np.abs(a[f["rl20_idx"]] - f["rl20_amps"]).max() <= 1e-15
    assert abs(a.sum() - f["rl20_sum"]) <= 1e-12
    assert abs(np.sum(np.abs(a) ** 2 * np.arange(2 ** n)) - f["rl20_wsum"]) <= 1e-6


# ---- 2. doctest known answers (state_vector.hpp:1396-1739) ----------------------
def basis(n, i):
    a = np.zeros(2 ** n, complex); a[i] = 1
    return a


def test_doctest_known_answers(orc):
    # CX, control 1 target 0: |10> -> |11> (index 2 -> 3) (:1557-1563)
    for i, j in [(0, 0), (1, 1), (2, 3), (3, 2)]:
        a = basis(2, i); orc.apply_gate(a, Gate("X", (0,), (1,)))
        assert np.array_equal(a, basis(2, j))
    # SWAP on targets {0, 1}: index 1 <-> 2 (:1519-1533)
    for i, j in [(0, 0), (1, 2), (2, 1), (3, 3)]:
        a = basis(2, i); orc.apply_gate(a, Gate("SWAP", (0, 1)))
        assert np.array_equal(a, basis(2, j))
    # CY == S . CX . SDG and CZ == H . CX . H for both orientations (:1576-1640)
    for c, t in [(0, 1), (1, 0)]:
        for i in range(4):
            a = basis(2, i); orc.apply_gate(a, Gate("Y", (t,), (c,)))
            b = basis(2, i); orc.apply_gates(b, [Gate("SDG", (t,)), Gate("X", (t,), (c,)), Gate("S", (t,))])
            assert np.abs(a - b).max() <= 1e-10
            a = basis(2, i); orc.apply_gate(a, Gate("Z", (t,), (c,)))
            b = basis(2, i); orc.apply_gates(b, [Gate("H", (t,)), Gate("X", (t,), (c,)), Gate("H", (t,))])
            assert np.abs(a - b).max() <= 1e-10
    # CH, control 1 target 0 (:1642-1665)
    r = 1 / math.sqrt(2)
    for i, want in [(0, [1, 0, 0, 0]), (1, [0, 1, 0, 0]), (2, [0, 0, r, r]), (3, [0, 0, r, -r])]:
        a = basis(2, i); orc.apply_gate(a, Gate("H", (0,), (1,)))
        assert np.abs(a - np.array(want)).max() <= 1e-10
    # I and TDG.T leave |1> alone (:1462-1495)
    a = basis(1, 1); orc.apply_gates(a, [Gate("I", (0,)), Gate("TDG", (0,)), Gate("T", (0,))])
    assert np.abs(a - basis(1, 1)).max() <= 1e-10
    # probabilities after X, Y, Z, H (:1667-1737)
    a = basis(1, 0)
    for g, p0 in [("X", 0), ("X", 1), ("Y", 0), ("Y", 1), ("Z", 1), ("Z", 1), ("H", 0.5), ("H", 1)]:
        orc.apply_gate(a, Gate(g, (0,)))
        assert abs(orc.probability(a, 0, 0) - p0) <= 1e-10
        assert abs(orc.probability(a, 0, 1) - (1 - p0)) <= 1e-10


# ---- 3. brute force with explicit operators -----------------------------------
def frac_angle(num, power):
    return -math.pi * num / 2 ** (power - 1)


def rz(t):
    return np.diag([np.exp(-0.5j * t), np.exp(0.5j * t)])


def rx(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return np.array([[c, -1j * s], [-1j * s, c]])


def ry(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return np.array([[c, -s], [s, c]])


def ri(t):
    return np.exp(-0.5j * t) * I2


def matrix_of(g: Gate) -> np.ndarray:
    """The operator each reference gate applies, written from its math (for
    two targets the index is 2*bit(targets[0]) + bit(targets[1]))."""
    name = g.name.upper()
    p = [complex(x).real for x in g.params] if not name.endswith("FRAC") else list(g.params)
    if name == "I": return I2
    if name == "X": return np.array([[0, 1], [1, 0]])
    if name == "Y": return np.array([[0, -1j], [1j, 0]])
    if name == "Z": return np.diag([1, -1])
    if name == "H": return HM
    if name == "S": return np.diag([1, 1j])
    if name == "T": return np.diag([1, np.exp(0.25j * math.pi)])
    if name == "SDG": return np.diag([1, -1j])
    if name == "TDG": return np.diag([1, np.exp(-0.25j * math.pi)])
    if name == "RX": return rx(p[0])
    if name == "RY": return ry(p[0])
    if name in ("RZ", "U1"): return rz(p[0])              # U1 is RZ in the reference
    if name == "RI": return ri(p[0])
    if name == "R1": return np.diag([1, np.exp(1j * p[0])])
    if name == "RXFRAC": return rx(frac_angle(*p))
    if name == "RYFRAC": return ry(frac_angle(*p))
    if name == "RZFRAC": return rz(frac_angle(*p))
    if name == "RIFRAC": return ri(frac_angle(*p))
    if name == "R1FRAC":                                     # RZ(-num) then RI(num), unsigned -num
        num, pw = p
        return ri(frac_angle(num, pw + 1)) @ rz(frac_angle((2 ** 32 - num) % 2 ** 32, pw + 1))
    if name == "U2":
        ph, la = p
        return np.array([[1, -np.exp(1j * la)], [np.exp(1j * ph), np.exp(1j * (ph + la))]]) / math.sqrt(2)
    if name == "U3":                                         # e^{-i(phi+lambda)/2} x textbook U3
        th, ph, la = p
        c, s = math.cos(th / 2), math.sin(th / 2)
        u = np.array([[c, -np.exp(1j * la) * s], [np.exp(1j * ph) * s, np.exp(1j * (ph + la)) * c]])
        return np.exp(-0.5j * (ph + la)) * u
    if name == "C1": return np.array(g.params, complex).reshape(2, 2)
    if name == "C2":                                         # the constructor repeats column 0
        c = np.array(g.params, complex).reshape(4, 4)
        return np.repeat(c[:, :1], 4, axis=1)
    if name == "SWAP": return np.eye(4)[[0, 2, 1, 3]]
    cnot = np.eye(4)[[0, 1, 3, 2]]                           # control targets[0], target targets[1]
    if name == "RZZ": return np.diag(np.exp(-0.5j * p[0] * np.array([1, -1, -1, 1])))
    if name == "RYY":
        a = np.kron(rx(math.pi / 2), rx(math.pi / 2))
        b = np.kron(rx(-math.pi / 2), rx(-math.pi / 2))
        return b @ cnot @ np.kron(I2, rz(p[0])) @ cnot @ a
    if name == "RXX":                                        # extra H on targets[1] where bit(targets[0]) = 0
        extra = np.kron(np.diag([1, 0]), HM) + np.kron(np.diag([0, 1]), I2)
        hh = np.kron(HM, HM)
        return hh @ cnot @ np.kron(I2, rz(p[0])) @ cnot @ extra @ hh
    raise KeyError(name)


def full_operator(g: Gate, n: int) -> np.ndarray:
    m = matrix_of(g)
    N = 2 ** n
    op = np.zeros((N, N), complex)
    ts = list(g.targets)
    cmask = sum(1 << c for c in g.controls)
    for i in range(N):
        if (i & cmask) != cmask:
            op[i, i] = 1
            continue
        loc = 0
        for t in ts:
            loc = loc * 2 + ((i >> t) & 1)
        base = i
        for t in ts:
            base &= ~(1 << t)
        for r in range(len(m)):
            j = base
            for k, t in enumerate(ts):
                if (r >> (len(ts) - 1 - k)) & 1:
                    j |= 1 << t
            op[j, i] += m[r, loc]
    return op


@pytest.mark.parametrize("name", ALL_ONE + ALL_TWO)
def test_brute_force_operators(orc, name):
    n = 5
    rng = np.random.default_rng(hash(name) % 2 ** 32)
    for trial in range(6):
        nt = 2 if name in ALL_TWO else 1
        nc = trial % 3
        qs = rng.choice(n, size=nt + nc, replace=False)
        g = Gate(name, tuple(int(q) for q in qs[:nt]), tuple(int(q) for q in qs[nt:]), params_for(name, rng))
        a = rand_state(n, trial)
        want = full_operator(g, n) @ a
        got = a.copy()
        orc.apply_gate(got, g)
        assert np.abs(got - want).max() <= 1e-13, g


def test_textbook_identities():
    # RZZ and RYY are the textbook exp(-i theta/2 P x P); RXX is not (its extra H)
    t = 0.731
    Z = np.diag([1, -1]); Y = np.array([[0, -1j], [1j, 0]]); X = np.array([[0, 1], [1, 0]])
    from scipy.linalg import expm
    assert np.abs(matrix_of(Gate("RZZ", (0, 1), (), (t,))) - expm(-0.5j * t * np.kron(Z, Z))).max() < 1e-12
    assert np.abs(matrix_of(Gate("RYY", (0, 1), (), (t,))) - expm(-0.5j * t * np.kron(Y, Y))).max() < 1e-12
    assert np.abs(matrix_of(Gate("RXX", (0, 1), (), (t,))) - expm(-0.5j * t * np.kron(X, X))).max() > 0.1


# ---- 4. closed forms ------------------------------------------------------------
def test_bv_closed_form(orc):
    n = 14
    gates, mask = bernstein_vazirani(n, seed=9)
    a = basis(n, 0)
    orc.apply_gates(a, gates)
    want = np.zeros(2 ** n, complex)
    want[mask] = 1 / math.sqrt(2)
    want[mask + 2 ** (n - 1)] = -1 / math.sqrt(2)
    assert np.abs(a - want).max() <= 1e-14


def test_qft_closed_form(orc):
    n = 11
    x = rand_state(n, 2)
    y = x.copy()
    orc.apply_gates(y, qft(n))
    rev = bit_reverse_perm(n)
    want = np.empty_like(x)
    want[rev] = np.fft.ifft(x[rev]) * math.sqrt(2 ** n)
    assert np.abs(y - want).max() <= 1e-13


# ---- 5. invariants --------------------------------------------------------------
def test_part_path_equals_per_gate(orc):
    from hisvsim_amd.partition import nat_parts
    n = 13
    gates = random_layered(n, 8, seed=4)
    for k in (3, 5, 8):
        a = rand_state(n, k)
        b, c = a.copy(), a.copy()
        orc.apply_gates(a, gates)
        # parts that keep the gate order: the same operations on every amplitude
        for p in nat_parts(gates, k):
            orc.run_part(b, p, k % 2)
        assert np.array_equal(a, b)
        # the greedy partitioner reorders commuting gates: equal up to rounding
        for p in greedy_parts(gates, k):
            orc.run_part(c, p, k % 2)
        assert np.abs(a - c).max() <= 1e-14


def test_permutation_round_trip(orc):
    n = 10
    rng = np.random.default_rng(0)
    perm = [int(x) for x in rng.permutation(n)]
    inv = [0] * n
    for s, d in enumerate(perm):
        inv[d] = s
    a = rand_state(n, 1)
    b = orc.permute(orc.permute(a, perm), inv)
    assert np.array_equal(a, b)
    # bit s of the index moves to bit perm[s]
    i = 0b1010011010
    e = basis(n, i)
    j = sum(((i >> s) & 1) << perm[s] for s in range(n))
    assert orc.permute(e, perm)[j] == 1


def test_exchange_order(orc):
    # svsim-mpi.hpp:139-149: kept qubits in slot order, then the moved ones in slot order
    assert orc.exchange_order([0, 1, 2, 3, 4, 5], [4, 1]) == [0, 2, 3, 5, 1, 4]
    assert orc.exchange_order([3, 0, 5, 1, 4, 2], [5, 2]) == [3, 0, 1, 4, 5, 2]


def test_measure_walk(orc):
    a = np.zeros(8, complex)
    a[1] = math.sqrt(0.25); a[6] = math.sqrt(0.75)
    u = np.array([0.1, 0.3, 0.99])
    out = orc.measure_sorted(a, [0, 1, 2], u)
    assert out.tolist() == [[1, 0, 0], [0, 1, 1], [0, 1, 1]]
