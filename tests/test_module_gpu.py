"""The drop-in boundary on the device: the svsimulator_py_wrapper module
driven the way the reference's driver drives its own module
(qasm_assembler_standalone.py:1778-2014 over py_wrapper.cpp:101-175):
StateVector(n, 0); gates built as AggregateGate_* -> AggregateGateFactory ->
BoundGate; applied one by one with BoundGate.apply or part by part with
gather_and_execute_on under set_opt_slots(1); read back through ==,
state_equal, measure and the amplitudes() extension. Every result is
checked against the CPU oracle (or, where the oracle would take minutes, the
circuit's closed form and the module's own per-gate path). The caller here
is this test's own code."""
import os

import numpy as np
import pytest

from gatesets import ALL_ONE, ALL_TWO, params_for, random_parts
from hisvsim_amd import qasm
from hisvsim_amd.circuit import Gate, random_layered
from test_qasm import QDIR, bv_mask, closed_form, phase_aligned_diff, prog_of

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def m():
    import hisvsim_amd
    mod = hisvsim_amd.load_module()
    yield mod
    mod.set_opt_slots(0)


def bound(m, g: Gate):
    """A BoundGate for a gate record, through the aggregate the reference's
    driver would pick for its parameter count (basic_gates.hpp:1527-1776)."""
    p = g.params
    if g.name.upper().endswith("FRAC"):
        ag = m.AggregateGate_Frac(g.name, int(p[0]), int(p[1]))
    elif len(p) == 0:
        ag = m.AggregateGate_0(g.name)
    elif len(p) == 1:
        ag = m.AggregateGate_1(g.name, complex(p[0]))
    elif len(p) == 2:
        ag = m.AggregateGate_2(g.name, *map(complex, p))
    elif len(p) == 3:
        ag = m.AggregateGate_3(g.name, *map(complex, p))
    elif len(p) == 4:
        ag = m.AggregateGate_C1(g.name, *map(complex, p))
    else:
        ag = m.AggregateGate_C2(g.name, *map(complex, p))
    return m.BoundGate(m.AggregateGateFactory(ag), list(g.controls), list(g.targets))


def dense_prep(n):
    """H then RY on every qubit: a dense start that the reference's caller can
    make from StateVector(n, 0) with gates alone."""
    return [Gate("H", (q,)) for q in range(n)] + [Gate("RY", (q,), (), (0.3 + 0.17 * q,)) for q in range(n)]


def zero_state(n):
    a = np.zeros(2 ** n, complex)
    a[0] = 1
    return a


def gate_cases():
    rng = np.random.default_rng(41)
    cases = []
    for name in ALL_ONE:
        for t, ctl in (((0,), ()), ((5,), (2,)), ((9,), (1, 7))):
            cases.append(Gate(name, t, ctl, params_for(name, rng)))
    for name in ALL_TWO:
        for t, ctl in (((0, 1), ()), ((8, 3), (5,)), ((2, 9), (0, 6))):
            cases.append(Gate(name, t, ctl, params_for(name, rng)))
    return cases


@pytest.mark.parametrize("g", gate_cases(), ids=lambda g: f"{g.name}-t{'.'.join(map(str, g.targets))}"
                                                        f"-c{'.'.join(map(str, g.controls))}")
def test_every_gate_type_through_the_aggregates(m, orc, g):
    """All 28 gate types (I, C1 and C2 included) with 0-2 controls, built
    through the aggregates and applied with BoundGate.apply on a dense
    10-qubit state, against the oracle to 1e-12."""
    n = 10
    gates = dense_prep(n) + [g]
    want = zero_state(n)
    orc.apply_gates(want, gates)
    sv = m.StateVector(n, 0)
    for x in gates:
        bound(m, x).apply(sv)
    got = np.asarray(sv.amplitudes())
    assert np.abs(got - want).max() <= TOL * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("seed", range(3))
def test_gather_and_execute_on_parts_with_fast_slots(m, orc, seed):
    """set_opt_slots(1) and gather_and_execute_on per part (every gate type
    except I, which the reference's execute_on aborts on), against the
    oracle's part driver with the same fast slots (execute.hpp:184-199,
    318-470), 14 qubits, 1e-12."""
    n = 14
    rng = np.random.default_rng(seed)
    names = [x for x in ALL_ONE + ALL_TWO if x not in ("I", "C1", "C2")]
    parts = random_parts(rng, n, 8, 6, 12, names=names, max_controls=1)
    want = zero_state(n)
    orc.apply_gates(want, dense_prep(n))
    for p in parts:
        orc.run_part(want, p, 1)
    m.set_opt_slots(1)
    try:
        sv = m.StateVector(n, 0)
        for x in dense_prep(n):
            bound(m, x).apply(sv)
        for p in parts:
            m.gather_and_execute_on(sv, [bound(m, g) for g in p])
    finally:
        m.set_opt_slots(0)
    got = np.asarray(sv.amplitudes())
    assert np.abs(got - want).max() <= TOL


def shipped_parts(name, partfile):
    prog = prog_of(name)
    return prog, qasm.parts_from_parts_file(prog, os.path.join(QDIR, partfile))


def test_bv19_shipped_smart_partition(m, orc):
    """bv_n19 through the reference's shipped dagP partition
    (bv_n19_part_smart) with fast slots: the module's part path equals the
    oracle's per-gate run and the closed form, == and state_equal agree with
    the per-gate BoundGate.apply run, and a seeded measure of the data
    register returns the hidden string in every shot."""
    prog, parts = shipped_parts("bv_n19", "bv_n19_part_smart")
    n = prog.num_qubits
    gates = prog.gates()
    want = zero_state(n)
    orc.apply_gates(want, gates)
    m.set_opt_slots(1)
    try:
        sv = m.StateVector(n, 0)
        for p in parts:
            m.gather_and_execute_on(sv, [bound(m, g) for g in p])
    finally:
        m.set_opt_slots(0)
    got = np.asarray(sv.amplitudes())
    assert np.abs(got - want).max() <= TOL
    d, dn = phase_aligned_diff(got, closed_form("bv_n19", prog))
    assert d <= TOL and dn <= TOL
    per_gate = m.StateVector(n, 0)
    for g in gates:
        bound(m, g).apply(per_gate)
    assert sv == per_gate and sv.state_equal(per_gate) and not (sv != per_gate)
    data = list(range(n - 1))
    shots = sv.measure(data, 16, seed=7)
    assert shots == sv.measure(data, 16, seed=7)
    mask = bv_mask(prog)
    assert all(shot == [(mask >> q) & 1 for q in data] for shot in shots)


def test_qft26_shipped_natural_partition(m):
    """qft_n26 through the shipped natural-order partition with fast slots:
    every amplitude has the closed-form magnitude 2^-13 (QFT of |0>), and the
    state equals the module's own per-gate run (== within 1e-10, and
    max |d| <= 1e-12 relative)."""
    prog, parts = shipped_parts("qft_n26", "qft_n26_part_nat")
    n = prog.num_qubits
    m.set_opt_slots(1)
    try:
        sv = m.StateVector(n, 0)
        for p in parts:
            m.gather_and_execute_on(sv, [bound(m, g) for g in p])
    finally:
        m.set_opt_slots(0)
    got = np.asarray(sv.amplitudes())
    assert np.abs(np.abs(got) - 2 ** (-n / 2)).max() <= TOL * 2 ** (-n / 2) * 16
    per_gate = m.StateVector(n, 0)
    for g in prog.gates():
        bound(m, g).apply(per_gate)
    assert sv == per_gate
    assert sv.max_abs_diff(per_gate) <= TOL * 2 ** (-n / 2)


def test_equality_detects_a_changed_gate(m):
    """== / != / state_equal (state_vector.hpp:1003-1072, threshold 1e-10):
    an RZ(1e-12) changes amplitudes by < 1e-12 and stays equal; an RZ(1e-3)
    changes them by ~1e-5 and is unequal."""
    n = 12
    a, b, c = m.StateVector(n, 0), m.StateVector(n, 0), m.StateVector(n, 0)
    for x in dense_prep(n):
        for s in (a, b, c):
            bound(m, x).apply(s)
    bound(m, Gate("RZ", (3,), (), (1e-12,))).apply(b)
    bound(m, Gate("RZ", (3,), (), (1e-3,))).apply(c)
    assert a == b and a.state_equal(b)
    assert a != c and not a.state_equal(c) and not (a == c)


def test_measure_is_seeded_and_follows_the_probabilities(m):
    """measure (state_vector.hpp:925-981) with a seed: the same shots twice;
    on a GHZ state every shot is all zeros or all ones, both seen."""
    n = 16
    sv = m.StateVector(n, 0)
    bound(m, Gate("H", (0,))).apply(sv)
    for q in range(1, n):
        bound(m, Gate("X", (q,), (q - 1,))).apply(sv)
    qubits = [0, 5, 9, 15]
    shots = sv.measure(qubits, 64, seed=3)
    assert shots == sv.measure(qubits, 64, seed=3)
    assert all(s in ([0] * 4, [1] * 4) for s in shots)
    assert {tuple(s) for s in shots} == {(0,) * 4, (1,) * 4}
    assert abs(sv.probability(7, 1) - 0.5) <= TOL


def test_execute_fused_chain_from_the_module(m, orc):
    """The bench's default path through the module: plan_chain (13-qubit
    parts, moving low slots, beam search) and execute_fused_chain on a
    20-qubit random layered circuit with I gates in it, against the oracle:
    infidelity <= 1e-6, unit norm."""
    n = 20
    gates = random_layered(n, 8, seed=12)
    gates = [Gate("I", (9,))] + gates[:40] + [Gate("I", (17,))] + gates[40:]
    bgs = [bound(m, g) for g in gates]
    part_of, lows = m.plan_chain(bgs, 13, 4, 2)
    parts = [[] for _ in lows]
    for b, p in zip(bgs, part_of):
        parts[p].append(b)
    want = zero_state(n)
    orc.apply_gates(want, gates)
    sv = m.StateVector(n, 0)
    m.execute_fused_chain(sv, parts, lows)
    got = np.asarray(sv.amplitudes())
    assert abs(np.vdot(got, got).real - 1) <= TOL
    assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-6
