"""Host-side checks of the product (CPU only, no GPU needed).

* libhsv.so loads and exports every function include/hsv.h declares;
* each gate's host lowering (the matrix the kernels apply) equals the operator
  the reference applies (the explicit matrices of test_oracle_golden, which
  the reference's golden vectors pin);
* the fused unitary equals the ordered product of its gates' operators, is
  unitary, and non-unitary parts are detected;
* the svsimulator_py_wrapper module exposes the reference's API and rejects
  what the reference would abort on;
* the partitioners produce k-bounded, order-preserving partitions.
"""
import math
import os
import re

import numpy as np
import pytest

from conftest import ROOT, rand_state
from gatesets import ALL_ONE, ALL_TWO, params_for, random_circuit
from hisvsim_amd import capi
from hisvsim_amd.circuit import Gate, random_layered
from hisvsim_amd.partition import (chain_parts, check_parts, closure_parts, dfs_order, dfs_parts, gate_dag,
                                   greedy_parts, layered_parts,
                                   merge_parts, nat_parts)
from test_oracle_golden import full_operator, matrix_of


def header_functions():
    text = open(os.path.join(ROOT, "include", "hsv.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hsv_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = capi.lib()
    names = header_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(capi.SIGNATURES), "capi signatures out of sync with hsv.h"


@pytest.mark.parametrize("name", ALL_ONE + ALL_TWO)
def test_lowered_matrix_matches_reference_operator(name):
    rng = np.random.default_rng(abs(hash(name)) % 2 ** 32)
    for _ in range(4):
        nt = 2 if name in ALL_TWO else 1
        g = Gate(name, tuple(range(nt)), (), params_for(name, rng))
        got = capi.gate_matrix(g)
        want = matrix_of(g)
        assert got.shape == want.shape
        assert np.abs(got - want).max() <= 1e-15, name


def test_lowering_rejects_bad_gates():
    with pytest.raises(capi.HsvError, match="GATE"):
        capi.gate_matrix(Gate("SWAP", (0,)))
    with pytest.raises(capi.HsvError, match="ARG"):
        capi.gate_matrix(Gate("X", (1,), (1,)))
    with pytest.raises(capi.HsvError, match="ARG"):
        capi.gate_matrix(Gate("H", (70,)))
    with pytest.raises(ValueError):
        capi.to_hsv_gate(Gate("NOTAGATE", (0,)))


@pytest.mark.parametrize("seed", range(4))
def test_fused_unitary_is_product_of_gate_operators(seed):
    rng = np.random.default_rng(seed)
    n = 9
    k = 3 + seed
    qs = sorted(int(q) for q in rng.choice(n, size=k, replace=False))
    gates = []
    for g in random_circuit(rng, k, 12, max_controls=1):
        gates.append(Gate(g.name, tuple(qs[t] for t in g.targets), tuple(qs[c] for c in g.controls), g.params))
    q, U, err = capi.fused_unitary(n, gates)
    assert err <= 1e-13
    # the part's qubits, padded with the lowest free qubits up to 4
    used = sorted({x for g in gates for x in g.qubits()})
    assert set(used) <= set(q) and len(q) == max(4, len(used)) and list(q) == sorted(q)
    # apply U on the part's qubits of a full n-qubit state == gate-by-gate operators
    a = rand_state(n, seed)
    want = a.copy()
    for g in gates:
        want = full_operator(g, n) @ want
    got = np.zeros_like(a)
    D = len(U)
    for base in range(2 ** n):
        if any((base >> x) & 1 for x in q):
            continue
        idx = [base + sum(((r >> i) & 1) << x for i, x in enumerate(q)) for r in range(D)]
        got[idx] = U @ a[idx]
    assert np.abs(got - want).max() <= 1e-13


def test_fused_unitary_flags_c2():
    g = Gate("C2", (0, 1), (), tuple(complex(i + 1) for i in range(16)))
    _, _, err = capi.fused_unitary(6, [g])
    assert err > 1.0


def test_fused_unitary_width_limit():
    gates = [Gate("H", (q,)) for q in range(8)]
    with pytest.raises(capi.HsvError, match="LIMIT"):
        capi.fused_unitary(10, gates)


# ---- the reference's Python API --------------------------------------------------
def test_module_api_surface():
    import hisvsim_amd
    m = hisvsim_amd.load_module()
    for name in ["AggregateGate_0", "AggregateGate_1", "AggregateGate_2", "AggregateGate_3", "AggregateGate_C1",
                 "AggregateGate_C2", "AggregateGate_Frac", "AggregateGateFactory", "StateVector", "BoundGate",
                 "gather_and_execute_on", "obtain_apply_time", "obtain_gate_counter", "obtain_move_counter",
                 "obtain_gather_time", "set_num_threads", "set_init_num_threads", "set_opt_slots"]:
        assert hasattr(m, name), name
    for meth in ["qubits", "measure", "state_equal", "dump_state", "__eq__"]:
        assert hasattr(m.StateVector, meth)
    # construction works without a device
    f = m.AggregateGateFactory(m.AggregateGate_1("rx", 0.3))
    bg = m.BoundGate(f, [2], [0])
    assert bg.targets() == [0] and bg.controls() == [2] and bg.gate_type() == 10
    m.AggregateGateFactory(m.AggregateGate_0("H"))
    m.AggregateGateFactory(m.AggregateGate_3("U3", 0.1, 0.2, 0.3))
    m.AggregateGateFactory(m.AggregateGate_C2("C2", *([0.5] * 16)))
    m.AggregateGateFactory(m.AggregateGate_Frac("RZFrac", 1, 3))


def test_module_rejects_what_the_reference_aborts_on():
    import hisvsim_amd
    m = hisvsim_amd.load_module()
    with pytest.raises(ValueError):
        m.AggregateGate_0("NOPE")
    with pytest.raises(ValueError):
        m.AggregateGate_0("RX")          # RX takes a parameter
    with pytest.raises(ValueError):
        m.AggregateGate_1("H", 0.1)      # H takes none
    with pytest.raises(ValueError):
        m.BoundGate(m.AggregateGateFactory(m.AggregateGate_0("X")), [0, 1, 2, 3, 4, 5], [6])


# ---- partitioners -----------------------------------------------------------------
@pytest.mark.parametrize("k", [2, 4, 7, 10])
def test_partitioners_are_valid(k):
    rng = np.random.default_rng(k)
    circuits = [random_layered(14, 6, seed=k), random_circuit(rng, 12, 150, max_controls=1)]
    for gates in circuits:
        kk = max(k, max(len(g.qubits()) for g in gates))
        for parts in (nat_parts(gates, kk), greedy_parts(gates, kk), layered_parts(gates, 14, max(kk, 3)),
                      dfs_parts(gates, kk), merge_parts(greedy_parts(gates, kk), kk),
                      merge_parts(nat_parts(gates, kk), kk)):
            check_parts(gates, parts, max(kk, 3))
        if kk >= 5:
            for parts in (nat_parts(gates, kk, always=range(3)), dfs_parts(gates, kk, always=range(3))):
                check_parts(gates, parts, kk)
                assert all(len(set().union(*[set(g.qubits()) for g in p]) | {0, 1, 2}) <= kk for p in parts)


def test_dfs_order_is_topological_and_depth_first():
    """Every gate comes after its DAG predecessors; on a circuit of two
    independent chains, the first chain is followed to its end before the
    second starts."""
    rng = np.random.default_rng(5)
    gates = random_circuit(rng, 10, 200, max_controls=1)
    order = dfs_order(gates)
    assert sorted(order) == list(range(len(gates)))
    pos = {g: i for i, g in enumerate(order)}
    for i, succ in enumerate(gate_dag(gates)):
        assert all(pos[i] < pos[j] for j in succ)
    chains = [Gate("H", (0,)), Gate("H", (2,)), Gate("X", (1,), (0,)), Gate("X", (3,), (2,)), Gate("H", (1,)),
              Gate("H", (3,))]
    assert dfs_order(chains) == [0, 2, 4, 1, 3, 5]


def test_merge_parts_never_adds_parts():
    gates = random_layered(20, 10, seed=2)
    for k in (6, 9, 12):
        for parts in (nat_parts(gates, k), greedy_parts(gates, k)):
            merged = merge_parts(parts, k)
            assert len(merged) <= len(parts)
            check_parts(gates, merged, k)
    # neighbours that fit are joined
    a, b = [Gate("H", (0,))], [Gate("H", (1,))]
    assert merge_parts([a, b], 2) == [a + b] and merge_parts([a, b], 1) == [a, b]


def test_greedy_beats_natural_order_on_layered():
    gates = random_layered(24, 20, seed=1)
    assert len(greedy_parts(gates, 8)) < len(nat_parts(gates, 8)) / 1.5


@pytest.mark.parametrize("n,depth,k", [(16, 12, 12), (20, 10, 13), (24, 6, 13)])
def test_chain_parts_valid(n, depth, k):
    """Moving-layout parts: an acyclic k-bounded partition whose tiles (part
    qubits + lows) fit k, lows[0] = qubits 0-4, each later lows drawn from
    the previous tile; no more parts than with fixed low qubits."""
    gates = random_layered(n, depth, seed=n)
    parts, lows = chain_parts(gates, k)
    check_parts(gates, parts, k)
    assert sorted(lows[0]) == [0, 1, 2, 3, 4]
    for p, (part, low) in enumerate(zip(parts, lows)):
        assert len(set(low)) == 5
        tile = set(low).union(*[set(g.qubits()) for g in part])
        assert len(tile) <= k
        if p + 1 < len(parts):
            assert set(lows[p + 1]) <= tile
    assert len(parts) <= len(closure_parts(gates, k, always=range(5)))


@pytest.mark.parametrize("n,depth,k,lookahead", [(14, 8, 10, 7), (18, 10, 12, 13), (22, 8, 13, 9)])
def test_chain_parts_native_equals_python(n, depth, k, lookahead):
    """hsv_chain_parts (libhsv, host only) gives the Python method's parts and
    lows gate for gate."""
    gates = random_layered(n, depth, seed=3 * n)
    py = chain_parts(gates, k, lookahead=lookahead, native=False)
    nat = chain_parts(gates, k, lookahead=lookahead, native=True)
    assert [[id(g) for g in p] for p in py[0]] == [[id(g) for g in p] for p in nat[0]]
    assert py[1] == nat[1]


def identity_circuits():
    """Circuits with I gates (QASM 'id' and 'u0' lower to I). In the first,
    qubit 9's only gate in the first part is an I and its next real gate
    comes first after that part: a planner that counts the I offers qubit 9
    as a next low although the part's tile does not hold it. In the second,
    qubit 9 sees an I before any other gate and qubit 13 only ever sees I."""
    first = [Gate("I", (9,))]
    for _ in range(3):
        first += [Gate("H", (q,)) for q in range(9)]
        first += [Gate("Z", (5 + i,), (i,)) for i in range(4)]
    first += [Gate("Z", (10,), (9,)), Gate("Z", (11,), (10,)), Gate("H", (9,)), Gate("Z", (9,), (0,)),
              Gate("H", (11,))]
    second = [Gate("I", (9,)), Gate("I", (13,))] + [g for g in random_layered(14, 6, seed=11) if 13 not in g.qubits()]
    second.insert(len(second) // 2, Gate("I", (13,)))
    return [first, second]


@pytest.mark.parametrize("circuit", [0, 1])
@pytest.mark.parametrize("native,beam", [(False, (1, 1)), (True, (1, 1)), (True, (4, 2))])
def test_chain_parts_leave_identities_out_of_tiles(circuit, native, beam):
    """hsv_apply_fused_chain leaves I gates out of a part's tile, so the
    planner must too: every lows[p + 1] lies within part p's non-identity
    qubits plus lows[p], and no tile holds a qubit only an I touches."""
    gates = identity_circuits()[circuit]
    parts, lows = chain_parts(gates, 10, lookahead=10, native=native, beam=beam)
    check_parts(gates, parts, 10, identities_free=True)
    assert sum(len(p) for p in parts) == len(gates)
    for p, (part, low) in enumerate(zip(parts, lows)):
        tile = set(low).union(*[set(g.qubits()) for g in part if g.name != "I"])
        assert len(tile) <= 10
        if p + 1 < len(parts):
            assert set(lows[p + 1]) <= tile, (p, lows[p + 1], sorted(tile))
        assert 13 not in set(lows[p]) - set(range(5))


@pytest.mark.parametrize("n,depth,k,beam", [(16, 12, 12, (4, 2)), (20, 10, 13, (3, 3)), (14, 8, 10, (2, 4))])
def test_chain_parts_beam(n, depth, k, beam):
    """The beam search over low sets (hsv_chain_parts, width and branch > 1):
    a valid chain (acyclic k-bounded parts, lows[0] = qubits 0-4, each later
    lows drawn from the previous tile, the gates in circuit order within
    each part) with no more parts than the greedy chain."""
    gates = random_layered(n, depth, seed=5 * n + k)
    greedy, _ = chain_parts(gates, k)
    parts, lows = chain_parts(gates, k, beam=beam)
    check_parts(gates, parts, k)
    assert sum(len(p) for p in parts) == len(gates)
    assert len(parts) <= len(greedy)
    assert sorted(lows[0]) == [0, 1, 2, 3, 4]
    pos = {id(g): i for i, g in enumerate(gates)}
    for p, (part, low) in enumerate(zip(parts, lows)):
        assert [pos[id(g)] for g in part] == sorted(pos[id(g)] for g in part)
        tile = set(low).union(*[set(g.qubits()) for g in part])
        assert len(set(low)) == 5 and len(tile) <= k
        if p + 1 < len(parts):
            assert set(lows[p + 1]) <= tile
    with pytest.raises(ValueError):
        chain_parts(gates, k, beam=beam, native=False)


@pytest.mark.parametrize("seed", range(4))
def test_part_steps_are_a_valid_split(orc, seed):
    """hsv_part_steps, the fused steps of a wide BF16-split part: every
    step acts on at most 5 qubits, the steps keep each step's gates in
    circuit order, applying them one after the other equals the part (oracle,
    n = 14), and there are no more of them than the qubit-set closure makes."""
    rng = np.random.default_rng(seed)
    n = 14
    gates = random_layered(n, 6, seed=seed)
    parts, _ = chain_parts(gates, 12)
    part = max(parts, key=len)
    steps = capi.part_steps(part)
    assert sum(len(s) for s in steps) == len(part)
    pos = {id(g): i for i, g in enumerate(part)}
    for s in steps:
        assert len(set().union(*[set(g.qubits()) for g in s])) <= 5
        assert [pos[id(g)] for g in s] == sorted(pos[id(g)] for g in s)
    assert len(steps) <= len(closure_parts(part, 5))
    a = rand_state(n, int(rng.integers(1 << 30)))
    want, got = a.copy(), a.copy()
    orc.apply_gates(want, part)
    for s in steps:
        orc.apply_gates(got, s)
    assert np.abs(got - want).max() <= 1e-12


def test_part_steps_limits():
    assert capi.part_steps([]) == []
    assert len(capi.part_steps([Gate("H", (q,)) for q in range(3)])) == 1
    with pytest.raises(capi.HsvError, match="LIMIT"):
        capi.part_steps([Gate("X", (0,), (1, 2, 3, 4, 5))])   # 6 qubits: wider than a step
    with pytest.raises(capi.HsvError, match="LIMIT"):
        capi.part_steps([Gate("H", (q,)) for q in range(40)])   # C(40, 5) candidate steps


def test_chain_parts_apply_like_the_circuit(orc):
    """Applying the chain's parts in order equals the circuit (oracle, n=12)."""
    gates = random_layered(12, 8, seed=4)
    parts, _ = chain_parts(gates, 9)
    a = rand_state(12, 2)
    want, got = a.copy(), a.copy()
    orc.apply_gates(want, gates)
    for p in parts:
        orc.apply_gates(got, p)
    assert np.abs(got - want).max() <= 1e-12


def test_layered_inverse_undoes_the_circuit(orc):
    """The inverse the full-size round-trip test uses: circuit then inverse
    is the identity on a random state (oracle, n = 10, 1e-12)."""
    from conftest import layered_inverse
    gates = random_layered(10, 6, seed=9)
    a = rand_state(10, 3)
    got = a.copy()
    orc.apply_gates(got, gates + layered_inverse(gates))
    assert np.abs(got - a).max() <= 1e-12
    half = a.copy()
    orc.apply_gates(half, gates)
    assert np.abs(half - a).max() > 0.1   # the circuit itself is far from the identity


def test_check_parts_detects_reordering():
    gates = [Gate("H", (0,)), Gate("X", (1,), (0,)), Gate("H", (0,))]
    with pytest.raises(ValueError):
        check_parts(gates, [[gates[2]], [gates[0], gates[1]]])


def test_bench_reads_the_measured_traffic_of_its_default_workloads():
    """bench.py's roofline.traffic comes from the committed PMC summary
    (profiles/traffic.json, read at run time, so the folder must travel with
    the tree): the default chain, the three-piece chain and the FP64 part
    path each have a row, within 1% of their algorithmic bytes."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    algo = 2 * 16 * 2 ** 33
    for kernel, mode, k in [("k_part_bf", "fusedbf16", 13), ("k_part_bf", "fusedbf16x3", 13), ("k_part_tile", "part", 12)]:
        t = bench.measured_traffic(kernel, 33, mode, k)
        assert t is not None and abs(t / algo - 1) <= 0.01, (kernel, mode, t)
    with open(os.path.join(ROOT, ".gpurunignore")) as f:
        ignored = [line.strip().rstrip("/") for line in f]
    assert "./profiles" not in ignored and "profiles" not in ignored
