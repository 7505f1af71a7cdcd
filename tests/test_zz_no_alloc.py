"""hsv_apply_fused_chain_noalloc: the chain entry point that allocates no
device memory (its scratch is a static array per GPU) against
hsv_apply_fused_chain, through the C ABI: bit-identical states on two calls
in a row (the second smaller, in the other dtype, so the shared array holds
the first call's contents when the second starts), both within 1e-6 of the
oracle; a part whose step data exceeds the array is refused before any pass
and leaves the state as it was."""
import numpy as np
import pytest

from conftest import rand_state
from hisvsim_amd import capi
from hisvsim_amd.circuit import Gate, random_layered
from hisvsim_amd.partition import chain_parts

pytestmark = pytest.mark.gpu


def run_chain(n, a, parts, lows, dtype, noalloc, mode="bf16split"):
    with capi.State(n, dtype) as s:
        s.write(a)
        s.apply_fused_chain(parts, lows, mode, renormalize=True, noalloc=noalloc)
        s.synchronize()
        return s.read()


@pytest.mark.parametrize("first,second", [((22, 10, 13, "complex128"), (16, 6, 12, "complex64")),
                                          ((20, 12, 14, "complex64"), (14, 8, 12, "complex128"))])
def test_noalloc_chain_equals_chain(orc, first, second):
    for n, depth, k, dtype in (first, second):
        gates = random_layered(n, depth, seed=n + depth)
        parts, lows = chain_parts(gates, k, beam=(4, 2))
        a = rand_state(n, depth)
        if dtype == "complex64":
            a = a.astype(np.complex64).astype(complex)
        want = a.copy()
        orc.apply_gates(want, gates)
        print(f"[noalloc] n={n} depth={depth} k={k} {dtype} parts={len(parts)}", flush=True)
        ref = run_chain(n, a, parts, lows, dtype, noalloc=False)
        got = run_chain(n, a, parts, lows, dtype, noalloc=True)
        assert np.array_equal(got, ref)
        assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-6
        assert abs(np.vdot(got, got).real - 1) <= 1e-6


@pytest.mark.parametrize("n,depth,k,dtype", [(20, 12, 13, "complex128"), (16, 8, 12, "complex64")])
def test_noalloc_chain_three_piece_split(orc, n, depth, k, dtype):
    """The three-piece split (HSV_FUSED_BF16X3, 12 KB of fragments per step)
    through the no-allocation entry point: bit-identical to the allocating
    one, within infidelity 1e-10 of the oracle."""
    gates = random_layered(n, depth, seed=2 * n + depth)
    parts, lows = chain_parts(gates, k, beam=(4, 2))
    a = rand_state(n, depth)
    if dtype == "complex64":
        a = a.astype(np.complex64).astype(complex)
    want = a.copy()
    orc.apply_gates(want, gates)
    ref = run_chain(n, a, parts, lows, dtype, noalloc=False, mode="bf16x3")
    got = run_chain(n, a, parts, lows, dtype, noalloc=True, mode="bf16x3")
    assert np.array_equal(got, ref)
    # both norms: the complex64 input is the FP32-rounded state (norm 1 - ~1e-9)
    assert 1 - abs(np.vdot(want, got)) ** 2 / (np.vdot(want, want).real * np.vdot(got, got).real) <= 1e-10


def test_noalloc_three_piece_refuses_beyond_19_steps():
    """12 KB of fragments per step: the array holds 19 steps of a part with
    the three-piece split (28 with the two-piece one), so a 13-qubit part of
    24 steps runs with two pieces and is refused with three."""
    n = 16
    part = many_step_part(gates=240, width=13)   # 24 steps
    steps = len(capi.part_steps(part))
    assert 19 < steps <= 28
    with capi.State(n) as s:
        s.write(rand_state(n, 4))
        s.apply_fused_chain([part], [[0, 1, 2, 3, 4]], "bf16split", renormalize=True, noalloc=True)
        with pytest.raises(capi.HsvError, match="no-allocation") as exc:
            s.apply_fused_chain([part], [[0, 1, 2, 3, 4]], "bf16x3", renormalize=True, noalloc=True)
        assert exc.value.code == -5


def many_step_part(seed=3, gates=400, width=14):
    """One part on qubits 0..width-1 of 16 whose gates cut into many fused
    5-qubit steps (with the defaults more than the no-allocation array holds,
    28)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(gates):
        if i % 2:
            a, b = (int(x) for x in rng.choice(width, 2, replace=False))
            out.append(Gate("Z", (a,), (b,)))
        else:
            out.append(Gate("U3", (int(rng.integers(width)),), (), tuple(float(x) for x in rng.uniform(0, 6.2, 3))))
    return out


def test_noalloc_refuses_a_part_beyond_its_scratch(orc):
    n = 16
    part = many_step_part()
    assert len(capi.part_steps(part)) > 28
    lows = [[0, 1, 2, 3, 4]]
    a = rand_state(n, 9)
    with capi.State(n) as s:
        s.write(a)
        with pytest.raises(capi.HsvError, match="no-allocation") as exc:
            s.apply_fused_chain([part], lows, renormalize=True, noalloc=True)
        assert exc.value.code == -5
        assert np.array_equal(s.read(), a)
    want = a.copy()
    orc.apply_gates(want, part)
    got = run_chain(n, a, [part], lows, "complex128", noalloc=False)
    assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-6
