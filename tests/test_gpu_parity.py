"""GPU parity: libhsv (through its C ABI) against the CPU oracle.

Bars: <= 1e-12 absolute per amplitude for the FP64 per-gate, part and fused
FP64 paths; infidelity <= 1e-6 for the FP32-split fused path; bit-exact for
permutations and measurement outcomes. Sizes are ones the oracle finishes in
seconds; the closed-form tests (BV, QFT) run at larger n.
"""
import math

import numpy as np
import pytest

from conftest import bit_reverse_perm, rand_state
from gatesets import (ALL_ONE, ALL_TWO, UNITARY_ONE, UNITARY_TWO, params_for, random_circuit,
                      random_parts)
from hisvsim_amd import capi
from hisvsim_amd.circuit import Gate, bernstein_vazirani, qft, random_layered

pytestmark = pytest.mark.gpu

TOL = 1e-12


def gpu_run(n, state, fn, dtype="complex128"):
    with capi.State(n, dtype) as s:
        s.write(state)
        fn(s)
        s.synchronize()
        return s.read()


def per_gate_cases():
    cases = []
    for name in ALL_ONE:
        for t in (0, 1, 2, 5, 9):
            for ctl in ((), (3,) if t != 3 else (4,), (8, 6) if t not in (8, 6) else (7, 4)):
                cases.append((name, (t,), ctl))
    for name in ALL_TWO:
        for t in ((0, 1), (1, 0), (2, 7), (9, 3), (4, 5)):
            for ctl in ((), tuple(q for q in (6, 8) if q not in t)[:1], (0, 8) if 0 not in t and 8 not in t else (6,)):
                cases.append((name, t, ctl))
    return cases


@pytest.mark.parametrize("name,targets,controls", per_gate_cases())
def test_per_gate_matches_oracle(orc, name, targets, controls):
    n = 10
    rng = np.random.default_rng(abs(hash((name, targets, controls))) % 2 ** 32)
    g = Gate(name, targets, controls, params_for(name, rng))
    a = rand_state(n, 1)
    want = a.copy()
    orc.apply_gate(want, g)
    got = gpu_run(n, a, lambda s: s.apply_gate(g))
    assert np.abs(got - want).max() <= TOL


@pytest.mark.parametrize("seed", range(4))
def test_random_circuit_per_gate(orc, seed):
    n = 12
    rng = np.random.default_rng(seed)
    gates = random_circuit(rng, n, 80, names=ALL_ONE + ALL_TWO, max_controls=3)
    a = rand_state(n, seed)
    want = a.copy()
    orc.apply_gates(want, gates)
    got = gpu_run(n, a, lambda s: s.apply_gates(gates))
    assert np.abs(got - want).max() <= TOL


@pytest.mark.parametrize("k,opt_slots,seed", [(3, 0, 0), (4, 1, 1), (6, 0, 2), (6, 1, 3), (8, 1, 4), (10, 0, 5), (11, 1, 6)])
def test_parts_match_oracle(orc, k, opt_slots, seed):
    n = 14
    rng = np.random.default_rng(100 + seed)
    parts = random_parts(rng, n, 6, k, 12, names=ALL_ONE + ALL_TWO, max_controls=2)
    a = rand_state(n, seed)
    want = a.copy()
    for p in parts:
        orc.run_part(want, p, opt_slots)
    per_gate = a.copy()
    for p in parts:
        orc.apply_gates(per_gate, p)
    assert np.array_equal(want, per_gate)  # the oracle's two paths agree bit for bit

    def run(s):
        for p in parts:
            s.apply_part(p, opt_slots)
    got = gpu_run(n, a, run)
    assert np.abs(got - want).max() <= TOL


def test_part_wider_than_lds_falls_back(orc):
    n = 18
    rng = np.random.default_rng(7)
    parts = random_parts(rng, n, 2, 16, 30)
    a = rand_state(n, 3)
    want = a.copy()
    for p in parts:
        orc.apply_gates(want, p)
    got = gpu_run(n, a, lambda s: [s.apply_part(p) for p in parts])
    assert np.abs(got - want).max() <= TOL


@pytest.mark.parametrize("k,seed", [(1, 0), (3, 1), (4, 2), (5, 3), (6, 4), (7, 5)])
def test_fused_f64_matches_oracle(orc, k, seed):
    n = 13
    rng = np.random.default_rng(200 + seed)
    parts = random_parts(rng, n, 4, k, 3 * k + 2, names=UNITARY_ONE + (UNITARY_TWO if k >= 2 else []),
                         max_controls=min(2, k - 1) if k > 1 else 0)
    a = rand_state(n, seed)
    want = a.copy()
    for p in parts:
        orc.apply_gates(want, p)
    got = gpu_run(n, a, lambda s: [s.apply_fused(p, "f64") for p in parts])
    assert np.abs(got - want).max() <= TOL


@pytest.mark.parametrize("k,seed", [(4, 0), (5, 1), (6, 2), (7, 3)])
def test_fused_f32split_infidelity(orc, k, seed):
    n = 14
    rng = np.random.default_rng(300 + seed)
    parts = random_parts(rng, n, 6, k, 3 * k, names=UNITARY_ONE + UNITARY_TWO, max_controls=1)
    a = rand_state(n, seed)
    want = a.copy()
    for p in parts:
        orc.apply_gates(want, p)
    got = gpu_run(n, a, lambda s: [s.apply_fused(p, "f32split") for p in parts])
    infid = 1 - abs(np.vdot(want, got)) ** 2
    assert infid <= 1e-6
    assert np.abs(got - want).max() <= 1e-6


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("k,seed,n", [(3, 4, 12), (5, 5, 14), (6, 6, 15), (7, 7, 16), (7, 8, 20)])
def test_fused_bf16split_infidelity(orc, k, seed, n, dtype):
    """BF16-split MFMA path (16 significant bits per operand, three products
    per real product, FP32 accumulation): infidelity and max error within
    1e-6 of the FP64 oracle over 6 random parts. Parts under 5 qubits are
    padded to 5."""
    rng = np.random.default_rng(700 + seed)
    parts = random_parts(rng, n, 6, k, 3 * k, names=UNITARY_ONE + UNITARY_TWO, max_controls=1)
    a = rand_state(n, seed)
    want = a.copy()
    for p in parts:
        orc.apply_gates(want, p)
    got = gpu_run(n, a, lambda s: [s.apply_fused(p, "bf16split") for p in parts] + [s.renormalize()], dtype)
    assert abs(np.vdot(got, got).real - 1) <= (1e-12 if dtype == "complex128" else 1e-6)
    infid = 1 - abs(np.vdot(want, got)) ** 2
    assert infid <= 1e-6
    assert np.abs(got - want).max() <= 1e-6


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("k,n,ngates,seed", [(8, 14, 40, 1), (10, 16, 60, 2), (12, 16, 90, 3), (12, 20, 400, 4),
                                             (13, 18, 120, 5), (13, 13, 60, 6), (14, 18, 150, 7), (14, 14, 80, 8)])
def test_fused_bf16split_wide_parts(orc, k, n, ngates, seed, dtype):
    """Parts of 8..14 qubits on the BF16-split path: split into fused steps
    of <= 5 qubits applied to a 12-, 13- or 14-qubit tile in LDS (the 400-gate
    part needs more steps than fit at once and runs in several passes);
    infidelity and max error within 1e-6 of the FP64 oracle."""
    rng = np.random.default_rng(900 + seed)
    parts = random_parts(rng, n, 3, k, ngates // 3, names=UNITARY_ONE + UNITARY_TWO, max_controls=2)
    a = rand_state(n, seed)
    want = a.copy()
    for p in parts:
        orc.apply_gates(want, p)
    got = gpu_run(n, a, lambda s: [s.apply_fused(p, "bf16split") for p in parts] + [s.renormalize()], dtype)
    infid = 1 - abs(np.vdot(want, got)) ** 2
    assert infid <= 1e-6
    assert np.abs(got - want).max() <= 1e-6


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("k,n,ngates,seed", [(3, 12, 12, 1), (5, 14, 30, 2), (7, 16, 60, 3), (8, 14, 40, 4),
                                             (12, 16, 90, 5), (12, 20, 400, 6), (10, 18, 80, 7), (13, 16, 120, 8),
                                             (13, 13, 60, 9), (13, 20, 500, 10)])
def test_fused_bf16x3_parts(orc, k, n, ngates, seed, dtype):
    """The three-piece BF16 split (HSV_FUSED_BF16X3: six products per real
    product, 24 significant bits per operand, FP32 accumulation): parts of
    3..13 qubits, each in a 12-qubit tile with slots 0-4 (fragments resident
    in LDS) or a 13-qubit one (fragments streamed; at most 8 steps a launch,
    so the 400- and 500-gate parts run in several passes). No
    renormalisation needed: infidelity within
    1e-11 and max error within 1e-6 relative of the FP64 oracle (FP32 state
    in LDS per pass; the two-piece split measures ~1e-9 on these), the norm
    within 1e-6."""
    rng = np.random.default_rng(1900 + seed)
    parts = random_parts(rng, n, 3, min(k, 7), ngates // 3, names=UNITARY_ONE + UNITARY_TWO, max_controls=2)
    if k > 7:
        lows = [q for q in range(5)]
        qs = lows + [int(q) for q in rng.choice(range(5, n), k - 5, replace=False)]
        parts = []
        for _ in range(3):
            part = []
            for _ in range(ngates // 3):
                g = random_circuit(rng, k, 1, names=UNITARY_ONE + UNITARY_TWO, max_controls=2)[0]
                part.append(Gate(g.name, tuple(qs[t] for t in g.targets), tuple(qs[c] for c in g.controls),
                                 g.params))
            parts.append(part)
    a = rand_state(n, seed)
    want = a.copy()
    for p in parts:
        orc.apply_gates(want, p)
    got = gpu_run(n, a, lambda s: [s.apply_fused(p, "bf16x3") for p in parts], dtype)
    scale = np.abs(want).max()
    infid = 1 - abs(np.vdot(want, got)) ** 2 / np.vdot(got, got).real
    print(f"bf16x3 k={k} n={n}: infidelity {infid:.2e}, max error {np.abs(got - want).max() / scale:.2e} relative",
          flush=True)
    assert infid <= 1e-11
    assert np.abs(got - want).max() <= 1e-6 * scale
    assert abs(np.vdot(got, got).real - 1) <= 1e-6


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("n,depth,k", [(16, 12, 12), (20, 14, 12), (18, 12, 13), (20, 14, 13)])
def test_fused_chain_bf16x3(orc, n, depth, k, dtype):
    """The moving-layout chain at k = 12 and 13 with the three-piece split:
    within infidelity 1e-11 of the FP64 oracle, and at least 100 times
    closer than the two-piece split on the same parts."""
    from hisvsim_amd.partition import chain_parts
    gates = random_layered(n, depth, seed=3 * n + depth)
    parts, lows = chain_parts(gates, k)
    a = rand_state(n, depth)
    want = a.copy()
    orc.apply_gates(want, gates)
    infid = {}
    for mode in ("bf16x3", "bf16split"):
        got = gpu_run(n, a, lambda s: [s.apply_fused_chain(parts, lows, mode, renormalize=True)], dtype)
        infid[mode] = max(1 - abs(np.vdot(want, got)) ** 2 / np.vdot(got, got).real, 1e-18)
        if mode == "bf16x3":
            assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max()
    print(f"chain n={n}: infidelity bf16x3 {infid['bf16x3']:.2e}, bf16split {infid['bf16split']:.2e}", flush=True)
    assert infid["bf16x3"] <= 1e-11
    assert infid["bf16x3"] * 100 <= infid["bf16split"]


def test_fused_bf16x3_refuses_parts_wider_than_its_tile():
    """A 10-qubit part off qubits 0-4 needs a 14-qubit tile: the three-piece
    split (12- and 13-qubit tiles) refuses it."""
    n = 16
    gates = [Gate("H", (q,)) for q in range(6, 16)]
    with capi.State(n) as s:
        with pytest.raises(capi.HsvError):
            s.apply_fused(gates, "bf16x3")


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
def test_fused_bf16split_wide_part_off_the_low_qubits(orc, dtype):
    """An 11-qubit part on qubits 2..12: the tile adds qubits 0 and 1 so that
    it holds qubits 0-4 (a 13-qubit tile), and the steps pad with them."""
    n = 16
    rng = np.random.default_rng(77)
    qs = list(range(2, 13))
    gates = []
    for i in range(60):
        if i % 3 == 2:
            c, t = rng.choice(qs, 2, replace=False)
            gates.append(Gate("X", (int(t),), (int(c),)))
        else:
            q = int(rng.choice(qs))
            gates.append(Gate("U3", (q,), (), tuple(float(x) for x in rng.uniform(0, 2 * np.pi, 3))))
    a = rand_state(n, 5)
    want = a.copy()
    orc.apply_gates(want, gates)
    got = gpu_run(n, a, lambda s: [s.apply_fused(gates, "bf16split"), s.renormalize()], dtype)
    assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-6
    assert np.abs(got - want).max() <= 1e-6


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("n,depth,k", [(14, 10, 12), (16, 12, 13), (20, 14, 13), (18, 8, 12), (20, 12, 14)])
def test_fused_chain_matches_oracle(orc, n, depth, k, dtype):
    """Parts with a moving qubit layout (hsv_apply_fused_chain): each pass
    also brings the next part's lows into slots 0-4, and one permutation
    restores the standard layout; within 1e-6 of the FP64 oracle."""
    from hisvsim_amd.partition import chain_parts
    gates = random_layered(n, depth, seed=n + depth)
    parts, lows = chain_parts(gates, k)
    assert any(sorted(low) != [0, 1, 2, 3, 4] for low in lows)   # the layout does move
    a = rand_state(n, depth)
    want = a.copy()
    orc.apply_gates(want, gates)
    got = gpu_run(n, a, lambda s: [s.apply_fused_chain(parts, lows), s.renormalize()], dtype)
    assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-6
    assert np.abs(got - want).max() <= 1e-6


@pytest.mark.parametrize("native", [False, True])
def test_fused_chain_with_identity_gates(orc, native):
    """I gates (QASM 'id', 'u0') anywhere in a chain, including on a qubit
    whose only gate in a part is the I and on a qubit that only I gates
    touch: the planner and hsv_apply_fused_chain both leave them out of the
    tiles, and the result matches the oracle within 1e-6."""
    from hisvsim_amd.partition import chain_parts
    from test_host import identity_circuits
    for gates in identity_circuits():
        n = 1 + max(q for g in gates for q in g.qubits())
        parts, lows = chain_parts(gates, 10 if n < 14 else 12, lookahead=10, native=native)
        a = rand_state(n, 5)
        want = a.copy()
        orc.apply_gates(want, gates)
        got = gpu_run(n, a, lambda s: [s.apply_fused_chain(parts, lows), s.renormalize()])
        assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-6


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("n,depth,k", [(14, 8, 12), (20, 12, 13), (22, 10, 13), (21, 9, 12)])
def test_fused_chain_renormalize(orc, n, depth, k, dtype):
    """HSV_CHAIN_RENORMALIZE: the chain ends on a unit-norm state (folded
    into the swap passes of the layout restore, or a separate renormalisation
    where the restore is one tile pass); equal to the chain followed by
    hsv_renormalize, and within 1e-6 of the FP64 oracle."""
    from hisvsim_amd.partition import chain_parts
    gates = random_layered(n, depth, seed=7 * n + k)
    parts, lows = chain_parts(gates, k)
    a = rand_state(n, k)
    want = a.copy()
    orc.apply_gates(want, gates)
    got = gpu_run(n, a, lambda s: s.apply_fused_chain(parts, lows, renormalize=True), dtype)
    two = gpu_run(n, a, lambda s: [s.apply_fused_chain(parts, lows), s.renormalize()], dtype)
    eps = 1e-12 if dtype == "complex128" else 1e-5
    assert abs(np.vdot(got, got).real - 1) <= eps
    assert np.abs(got - two).max() <= (1e-14 if dtype == "complex128" else 1e-6)
    assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-6
    assert np.abs(got - want).max() <= 1e-6


@pytest.mark.parametrize("form,tpw", [(1, 0), (1, 3), (2, 0), (2, 2)])
@pytest.mark.parametrize("n,depth,k", [(16, 12, 13), (20, 12, 13), (19, 30, 13)])
def test_fused_chain_part13_forms(orc, n, depth, k, form, tpw):
    """The other forms of the 13-qubit wide part kernel (part13_u): 1 = U
    fragments resident in LDS (1024 threads, next tile prefetched during the
    steps), 2 = fragments read by each wave from L2 into registers; with the
    default and a ragged number of tiles per workgroup: equal to the default
    LDS-streamed form within FP32 rounding and within 1e-6 of the FP64
    oracle."""
    from hisvsim_amd.partition import chain_parts
    gates = random_layered(n, depth, seed=11 * n + depth)
    parts, lows = chain_parts(gates, k)
    a = rand_state(n, depth)
    want = a.copy()
    orc.apply_gates(want, gates)

    def run(f):
        def fn(s):
            s.tune("part13_u", f)
            s.tune("part_tiles_per_wg", tpw)
            s.apply_fused_chain(parts, lows, renormalize=True)
        return gpu_run(n, a, fn)
    got, streamed = run(form), run(0)
    assert np.abs(got - streamed).max() <= 1e-6
    assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-6
    assert np.abs(got - want).max() <= 1e-6


@pytest.mark.parametrize("n,depth,beam", [(20, 14, (4, 4)), (22, 12, (16, 8))])
def test_fused_chain_beam_plan(orc, n, depth, beam):
    """The bench's plan: the chain planned with a beam over the low sets
    (hsv_chain_parts) and run from |0...0> with the folded renormalisation;
    within 1e-6 of the FP64 oracle, unit norm."""
    from hisvsim_amd.partition import chain_parts
    gates = random_layered(n, depth, seed=3 * n + depth)
    parts, lows = chain_parts(gates, 13, beam=beam)
    want = np.zeros(2 ** n, complex)
    want[0] = 1
    orc.apply_gates(want, gates)
    with capi.State(n) as s:
        s.apply_fused_chain(parts, lows, renormalize=True)
        got = s.read()
    assert abs(np.vdot(got, got).real - 1) <= 1e-12
    assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-6
    assert np.abs(got - want).max() <= 1e-6


@pytest.mark.parametrize("n,dtype", [(30, "complex128"), (31, "complex64"), (33, "complex128"), (34, "complex64")])
def test_fused_chain_round_trip_full_size(n, dtype):
    """A size-independent check at 17 GB and at 137 GB states (33 qubits
    complex128 and 34 complex64, the benchmark sizes): a random layered circuit
    followed by its exact inverse, planned as one beam chain and run from
    |0...0> with the folded renormalisation, must come back to |0...0>:
    |<0|psi>|^2 >= 1 - 1e-6 (the BF16-split bar) and unit norm."""
    from conftest import layered_inverse
    from hisvsim_amd.partition import chain_parts
    gates = random_layered(n, 12, seed=n)
    parts, lows = chain_parts(gates + layered_inverse(gates), 13, beam=(4, 4))
    with capi.State(n, dtype) as s:
        s.apply_fused_chain(parts, lows, renormalize=True)
        a0 = complex(s.read(0, 1)[0])
        norm2 = s.norm2()
    assert abs(a0) ** 2 >= 1 - 1e-6, (a0, n, dtype)
    assert abs(norm2 - 1) <= (1e-12 if dtype == "complex128" else 1e-5)


@pytest.mark.parametrize("n,dtype", [(33, "complex128"), (34, "complex64")])
def test_three_piece_chain_round_trip_full_size(n, dtype):
    """The same round trip on the three-piece split at the benchmark sizes
    (2^20 and 2^21 tiles of 13 qubits, so the kernel's column signs hash
    its largest tile indices): unit norm, and the probability mass left
    outside |0...0> (the FP64 norm of the state with amplitude 0 cleared,
    which complex64's rounding of a0 itself cannot hide) <= 1e-11
    (measured 2.4e-13 at 33 qubits complex128, 2.5e-13 at 34 complex64)."""
    from conftest import layered_inverse
    from hisvsim_amd.partition import chain_parts
    gates = random_layered(n, 12, seed=n)
    parts, lows = chain_parts(gates + layered_inverse(gates), 13, beam=(4, 4))
    with capi.State(n, dtype) as s:
        s.apply_fused_chain(parts, lows, "bf16x3", renormalize=True)
        norm2 = s.norm2()
        a0 = s.read(0, 1)
        s.write(np.zeros_like(a0), 0)
        rest = s.norm2()
    print(f"three-piece round trip n={n} {dtype}: mass outside |0> {rest:.2e}, "
          f"|norm2 - 1| = {abs(norm2 - 1):.2e}", flush=True)
    assert rest <= 1e-11, (rest, n, dtype)
    assert abs(norm2 - 1) <= (1e-12 if dtype == "complex128" else 1e-5)


def test_fused_chain_renormalize_zero_state():
    """A zero state stays zero (no error) under the renormalising chain."""
    from hisvsim_amd.partition import chain_parts
    n = 20
    gates = random_layered(n, 8, seed=5)
    parts, lows = chain_parts(gates, 13)
    got = gpu_run(n, np.zeros(2 ** n, complex), lambda s: s.apply_fused_chain(parts, lows, renormalize=True))
    assert not got.any()


@pytest.mark.parametrize("prep", ["reset", "reset+x", "write"])
def test_fused_chain_from_basis_state(orc, prep):
    """From |0...0> (after reset) the chain's first parts skip the tiles that
    are still zero; any other call in between (a gate, a write) must turn
    that off. Checked against the oracle from the same start."""
    from hisvsim_amd.partition import chain_parts
    n = 18
    gates = random_layered(n, 10, seed=21)
    parts, lows = chain_parts(gates, 13)
    top = n - 1   # the first part does not touch every qubit
    assert top not in {q for g in parts[0] for q in g.qubits()}
    want = np.zeros(2 ** n, complex)
    want[0] = 1
    if prep == "reset+x":
        orc.apply_gate(want, Gate("X", (top,)))
    orc.apply_gates(want, gates)
    with capi.State(n) as s:
        if prep == "write":
            s.write(np.eye(1, 2 ** n, 0, dtype=complex)[0])
        else:
            s.reset()
        if prep == "reset+x":
            s.apply_gate(Gate("X", (top,)))
        s.apply_fused_chain(parts, lows)
        s.renormalize()
        got = s.read()
    assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-6
    assert np.abs(got - want).max() <= 1e-6


def test_fused_chain_rejects_bad_lows(orc):
    """Argument errors are found before any pass: the state is unchanged."""
    from hisvsim_amd.partition import chain_parts
    n = 14
    gates = random_layered(n, 6, seed=3)
    parts, lows = chain_parts(gates, 12)
    a = rand_state(n, 1)
    with capi.State(n) as s:
        s.write(a)
        bad_first = [[1, 2, 3, 4, 5]] + lows[1:]
        with pytest.raises(capi.HsvError, match="ARG"):
            s.apply_fused_chain(parts, bad_first)
        tile0 = set(lows[0]).union(*[set(g.qubits()) for g in parts[0]])
        outside = [q for q in range(n) if q not in tile0]
        if len(parts) > 1 and outside:
            bad_next = [lows[0], [outside[0]] + lows[1][1:]] + lows[2:]
            with pytest.raises(capi.HsvError, match="ARG"):
                s.apply_fused_chain(parts, bad_next)
        with pytest.raises(capi.HsvError, match="ARG"):
            s.apply_fused_chain(parts, lows, "f64")
        assert np.array_equal(s.read(), a)


def test_fused_bf16split_wide_rejects_six_qubit_gate():
    g = Gate("SWAP", (0, 1), (2, 3, 4, 5))   # 6 qubits: wider than a 5-qubit step
    with capi.State(14) as s:
        with pytest.raises(capi.HsvError):
            s.apply_fused([g] + [Gate("H", (q,)) for q in range(6, 10)], "bf16split")


def test_fused_bf16split_limits():
    with capi.State(8) as s:   # 5-qubit part leaves 3 qubits outside: below the 4 the path needs
        with pytest.raises(capi.HsvError):
            s.apply_fused([Gate("H", (q,)) for q in range(5)], "bf16split")


def test_fused_rejects_non_unitary():
    g = Gate("C2", (0, 1), (), tuple(complex(i + 1, 0) for i in range(16)))
    with capi.State(8) as s:
        with pytest.raises(capi.HsvError, match="UNITARY"):
            s.apply_fused([g])


def permute_in_form(s, perm, form):
    """hsv_permute_qubits in a forced form: 1 = in place (two involution
    swap passes), 2 = out of place (one pass into a second buffer)."""
    s.tune("permute_in_place", form)
    s.permute_qubits(perm)


@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("seed", range(3))
def test_permute_matches_oracle(orc, seed, form):
    n = 13
    rng = np.random.default_rng(seed)
    perm = [int(x) for x in rng.permutation(n)]
    a = rand_state(n, seed)
    want = orc.permute(a, perm)
    got = gpu_run(n, a, lambda s: permute_in_form(s, perm, form))
    assert np.array_equal(got, want)


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("kind", ["cycle", "involution", "random"])
def test_permute_in_place_matches_oracle(orc, kind, dtype):
    """The in-place form bit for bit against the oracle at 20 qubits: one
    long cycle (two involutions), a set of disjoint swaps (one), and a
    random permutation; complex64 states are permuted exactly as well."""
    n = 20
    if kind == "cycle":
        perm = [(q + 7) % n for q in range(n)]
    elif kind == "involution":
        perm = list(range(n))
        for a_, b_ in ((0, 19), (3, 11), (4, 5), (8, 15)):
            perm[a_], perm[b_] = b_, a_
    else:
        perm = [int(x) for x in np.random.default_rng(9).permutation(n)]
    a = rand_state(n, 17)
    if dtype == "complex64":
        a = a.astype(np.complex64).astype(complex)
    want = orc.permute(a, perm)
    got = gpu_run(n, a, lambda s: permute_in_form(s, perm, 1), dtype)
    assert np.array_equal(got, want)


def test_permute_default_form_is_exact(orc):
    """Whatever form the default picks (out of place below 32 qubits when a
    second state fits), the result is the oracle's."""
    n = 16
    perm = [int(x) for x in np.random.default_rng(4).permutation(n)]
    a = rand_state(n, 2)
    assert np.array_equal(gpu_run(n, a, lambda s: s.permute_qubits(perm)), orc.permute(a, perm))


def test_exchange_order_permutation(orc):
    # the reference's qubit-exchange order (svsim-mpi.hpp:139-149) as a device permutation
    n = 12
    order = orc.exchange_order(list(range(n)), [2, 9, 5])
    new_slot = [0] * n
    for slot, q in enumerate(order):
        new_slot[q] = slot
    a = rand_state(n, 5)
    want = orc.permute(a, new_slot)
    got = gpu_run(n, a, lambda s: s.permute_qubits(new_slot))
    assert np.array_equal(got, want)


def test_reductions(orc):
    n = 15
    a = rand_state(n, 11) * 0.9
    b = rand_state(n, 12)
    with capi.State(n) as s, capi.State(n) as t:
        s.write(a)
        t.write(b)
        assert abs(s.norm2() - orc.norm2(a)) <= 1e-13
        ip = s.inner(t)
        assert abs(ip - np.vdot(a, b)) <= 1e-13
        assert abs(s.max_abs_diff(t) - np.abs(a - b).max()) <= 1e-15
        for q in (0, 7, 14):
            for v in (0, 1):
                assert abs(s.probability(q, v) - orc.probability(a, q, v)) <= 1e-12


def test_measure_matches_oracle(orc):
    n = 12
    a = rand_state(n, 21)
    u = np.sort(np.random.default_rng(3).uniform(0, 1, 200))
    qubits = [0, 5, 11, 3]
    want = orc.measure_sorted(a, qubits, u)
    with capi.State(n) as s:
        s.write(a)
        got = s.measure(qubits, u)
    assert np.array_equal(got, want)


def test_bv_closed_form():
    n = 24
    gates, mask = bernstein_vazirani(n, seed=5)
    with capi.State(n) as s:
        s.apply_gates(gates)
        amps = s.read()
    want = np.zeros(2 ** n, complex)
    want[mask] = 1 / math.sqrt(2)
    want[mask + 2 ** (n - 1)] = -1 / math.sqrt(2)
    assert np.abs(amps - want).max() <= TOL


def test_qft_closed_form():
    # y[rev(k)] = N^-1/2 sum_j x[rev(j)] e^{+2 pi i jk/N}
    n = 18
    x = rand_state(n, 4)
    rev = bit_reverse_perm(n)
    want = np.empty_like(x)
    want[rev] = np.fft.ifft(x[rev]) * math.sqrt(2 ** n)
    got = gpu_run(n, x, lambda s: s.apply_gates(qft(n)))
    err = np.abs(got - want).max()
    print(f"QFT18: max error {err:.2e}", flush=True)
    assert err <= TOL


@pytest.mark.parametrize("seed", range(2))
def test_complex64_state(orc, seed):
    n = 12
    rng = np.random.default_rng(seed)
    gates = random_circuit(rng, n, 40)
    a = rand_state(n, seed)
    want = a.copy()
    orc.apply_gates(want, gates)
    got = gpu_run(n, a.astype(np.complex64).astype(np.complex128), lambda s: s.apply_gates(gates), "complex64")
    assert np.abs(got - want).max() <= 1e-5
    parts = random_parts(rng, n, 4, 5, 8)
    want = a.copy()
    for p in parts:
        orc.apply_gates(want, p)
    got = gpu_run(n, a, lambda s: [s.apply_part(p) for p in parts], "complex64")
    assert np.abs(got - want).max() <= 1e-5
    got = gpu_run(n, a, lambda s: [s.apply_fused(p, "f32split") for p in parts], "complex64")
    assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-6


def test_modes_agree_at_26_qubits():
    """Size-independent properties at a larger n: the per-gate, LDS-part and
    fused-FP64 paths agree to 1e-12 and preserve the norm."""
    n = 26
    gates = random_layered(n, 2, seed=3)
    from hisvsim_amd.partition import layered_parts
    parts = layered_parts(gates, n, 6)
    with capi.State(n) as a, capi.State(n) as b, capi.State(n) as c:
        a.apply_gates(gates)
        for p in parts:
            b.apply_part(p)
        for p in parts:
            c.apply_fused(p, "f64")
        assert a.max_abs_diff(b) <= TOL
        assert a.max_abs_diff(c) <= TOL
        assert abs(a.norm2() - 1) <= 1e-12


def test_errors_are_reported():
    with capi.State(6) as s:
        with pytest.raises(capi.HsvError, match="ARG"):
            s.apply_gate(Gate("H", (6,)))
        with pytest.raises(capi.HsvError, match="ARG"):
            s.apply_gate(Gate("X", (1,), (1,)))
        with pytest.raises(capi.HsvError, match="GATE"):
            s.apply_gate(Gate("SWAP", (1,)))
        with pytest.raises(capi.HsvError, match="ARG"):
            s.permute_qubits([0, 0, 1, 2, 3, 4])


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("k,reg_bits", [(8, 0), (9, 2), (10, 3), (11, 4), (12, 0), (13, 4), (13, 5), (12, 5)])
def test_register_part_kernel(orc, dtype, k, reg_bits):
    """Parts of >= 8 qubits (with the low slots) run on the register-blocked
    kernel; every gate type, controls on register and thread bits."""
    n = 16
    rng = np.random.default_rng(1000 + 17 * k + reg_bits)
    names = [g for g in ALL_ONE + ALL_TWO if g != "I"]
    parts = random_parts(rng, n, 5, k, 4 * k, names=names, max_controls=2)
    a = rand_state(n, k)
    want = a.copy()
    for p in parts:
        orc.apply_gates(want, p)
    with capi.State(n, dtype) as s:
        s.tune("reg_bits", reg_bits)
        s.tune("low_slots", 0)
        s.write(a)
        for p in parts:
            s.apply_part(p, 0)
        got = s.read()
    # relative to the largest amplitude: C1/C2 are not unitary and grow the norm
    scale = max(1.0, np.abs(want).max())
    tol = (TOL if dtype == "complex128" else 1e-5) * scale
    assert np.abs(got - want).max() <= tol


def test_register_and_lds_kernels_agree():
    n = 20
    gates = random_layered(n, 6, seed=8)
    from hisvsim_amd.partition import greedy_parts
    parts = greedy_parts(gates, 9)
    with capi.State(n) as a, capi.State(n) as b:
        b.tune("part_kernel", 0)
        for p in parts:
            a.apply_part(p, 0)
            b.apply_part(p, 0)
        assert a.max_abs_diff(b) <= TOL


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("k,low", [(3, 0), (5, 5), (8, 5), (10, 5), (7, 0), (12, 0), (13, 0), (9, 3)])
def test_tile_part_kernel(orc, dtype, k, low):
    """part_kernel 2: parts on the FP64 tile kernel (the part's qubits with
    slots 0-4 padded to a 12-qubit tile for complex128, 13 for complex64;
    parts that do not fit take the register kernel); every gate type except
    I, 0-2 controls, the low slots added or not, against the oracle."""
    n = 17
    rng = np.random.default_rng(2000 + 31 * k + low)
    names = [g for g in ALL_ONE + ALL_TWO if g != "I"]
    parts = random_parts(rng, n, 6, k, 4 * k, names=names, max_controls=2)
    a = rand_state(n, k + low)
    want = a.copy()
    for p in parts:
        orc.apply_gates(want, p)
    with capi.State(n, dtype) as s:
        s.tune("part_kernel", 2)
        s.tune("low_slots", low)
        s.write(a)
        for p in parts:
            s.apply_part(p, 0)
        got = s.read()
    scale = max(1.0, np.abs(want).max())
    tol = (TOL if dtype == "complex128" else 1e-5) * scale
    assert np.abs(got - want).max() <= tol


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_tile_part_kernel_layered_fast_segments(orc, seed):
    """Random layered circuits (U3 on every qubit, CZ on random pairs, an
    extra Z, CCZ and controlled U3 now and then) on the tile kernel: the
    segments of uncontrolled one-qubit gates and sign flips run as rounds
    (dense gates by register slot, then one merged sign mask), the others
    through the general gate loop. Against the oracle at 1e-12."""
    n = 16
    rng = np.random.default_rng(77 + seed)
    gates = random_layered(n, 6, seed=5 + seed)
    extra = []
    for g in gates:
        extra.append(g)
        u = rng.random()
        if u < 0.05:
            extra.append(Gate("Z", (int(rng.integers(n)),)))
        elif u < 0.08:
            q = [int(x) for x in rng.choice(n, 3, replace=False)]
            extra.append(Gate("Z", (q[0],), (q[1], q[2])))
        elif u < 0.1:
            q = [int(x) for x in rng.choice(n, 2, replace=False)]
            extra.append(Gate("U3", (q[0],), (q[1],), (0.3, 0.7, 1.1)))
    from hisvsim_amd.partition import closure_parts
    parts = closure_parts(extra, 12, always=range(5))
    a = rand_state(n, 3)
    want = a.copy()
    orc.apply_gates(want, extra)
    with capi.State(n) as s:
        s.tune("part_kernel", 2)
        s.write(a)
        for p in parts:
            s.apply_part(p, 0)
        got = s.read()
    assert np.abs(got - want).max() <= TOL * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("dtype", ["complex128", "complex64"])
@pytest.mark.parametrize("n", [14, 21])
def test_tile_part_kernel_256_threads(orc, dtype, n):
    """tile_bits one below the default: 11-qubit tiles of 256 threads for
    complex128 (12 for complex64; 32 KB of LDS, four workgroups per CU) on
    layered circuits with extra CCZ and controlled U3 gates (rounds and the
    general gate loop), parts of up to that width with slots 0-4, against
    the oracle at 1e-12 (complex64 1e-5)."""
    rng = np.random.default_rng(n)
    gates = []
    for g in random_layered(n, 6, seed=n):
        gates.append(g)
        u = rng.random()
        if u < 0.04:
            q = [int(x) for x in rng.choice(n, 3, replace=False)]
            gates.append(Gate("Z", (q[0],), (q[1], q[2])))
        elif u < 0.08:
            q = [int(x) for x in rng.choice(n, 2, replace=False)]
            gates.append(Gate("U3", (q[0],), (q[1],), (0.3, 0.7, 1.1)))
    from hisvsim_amd.partition import closure_parts
    kt = 11 if dtype == "complex128" else 12
    parts = closure_parts(gates, kt, always=range(5))
    a = rand_state(n, 4)
    if dtype == "complex64":
        a = a.astype(np.complex64).astype(complex)
    print(f"[tile_bits] n={n} {dtype} kt={kt} parts={len(parts)}", flush=True)
    want = a.copy()
    orc.apply_gates(want, gates)
    with capi.State(n, dtype) as s:
        s.tune("part_kernel", 2)
        with pytest.raises(capi.HsvError):
            s.tune("tile_bits", kt - 1)   # only the default width and one less
        s.tune("tile_bits", kt)
        s.write(a)
        for p in parts:
            s.apply_part(p, 0)
        got = s.read()
    tol = TOL if dtype == "complex128" else 1e-5
    assert np.abs(got - want).max() <= tol * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("kernel", [1, 2])
def test_tile_and_register_kernels_agree_on_layered(kernel):
    """The tile and register part kernels against the LDS kernel on a
    random layered circuit's greedy parts (12 qubits with the low ones), at
    22 qubits, to 1e-12 relative to the largest amplitude."""
    from hisvsim_amd.partition import greedy_parts
    n = 22
    gates = random_layered(n, 8, seed=12)
    parts = greedy_parts(gates, 12, always=range(5))
    with capi.State(n) as a, capi.State(n) as b:
        a.tune("part_kernel", kernel)
        b.tune("part_kernel", 0)
        for p in parts:
            a.apply_part(p, 0)
        b.apply_gates(gates)
        assert a.max_abs_diff(b) <= TOL * b.max_abs()
