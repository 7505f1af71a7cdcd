"""The benchmark workloads at their full sizes, checked through properties
that hold at any size: closed forms where the circuit has one, agreement of
the FP64 paths to 1e-12, infidelity of the split-precision fused path within
1e-6, norm preservation. (The CPU oracle covers the same generators at sizes
it finishes in seconds, in the other test files.)"""
import math

import numpy as np
import pytest

from conftest import bit_reverse_perm
from hisvsim_amd import capi
from hisvsim_amd.circuit import Gate, bernstein_vazirani, gaussian_state, qft, random_layered
from hisvsim_amd.partition import chain_parts, check_parts, closure_parts

pytestmark = pytest.mark.gpu
TOL = 1e-12


def fused_infidelity(a: capi.State, b: capi.State) -> float:
    return 1.0 - abs(a.inner(b)) ** 2 / (a.norm2() * b.norm2())


def test_bv24_parts_of_8():
    """BV, 24 qubits (23 data + ancilla), seed-drawn mask, parts of at most 8
    qubits: per-gate, register/LDS part and BF16-split fused paths against
    amp[mask] = 1/sqrt2, amp[mask + 2^23] = -1/sqrt2."""
    n = 24
    gates, mask = bernstein_vazirani(n, seed=2024)
    parts = closure_parts(gates, 8)
    check_parts(gates, parts, 8)
    want = np.zeros(2 ** n, complex)
    want[mask], want[mask + 2 ** (n - 1)] = 1 / math.sqrt(2), -1 / math.sqrt(2)
    with capi.State(n) as s:
        s.apply_gates(gates)
        assert np.abs(s.read() - want).max() <= TOL
        s.reset()
        for p in parts:
            s.apply_part(p, 1)
        assert np.abs(s.read() - want).max() <= TOL
        s.reset()
        for p in closure_parts(gates, 7):
            s.apply_fused(p, "bf16split")
        s.renormalize()
        got = s.read()
    assert 1 - abs(np.vdot(want, got)) ** 2 <= 1e-6


def test_qft24_gaussian_closed_form():
    """QFT structure on a seeded complex Gaussian input: per-gate FP64 equals
    the bit-reversal-conjugated inverse DFT (numpy) to 1e-12."""
    n = 24
    x = gaussian_state(n, seed=7)
    rev = bit_reverse_perm(n)
    want = np.empty_like(x)
    want[rev] = np.fft.ifft(x[rev]) * math.sqrt(2 ** n)
    with capi.State(n) as s:
        s.write(x)
        s.apply_gates(qft(n))
        err = np.abs(s.read() - want).max()
        print(f"QFT24 Gaussian: max error {err:.2e}", flush=True)
        assert err <= TOL


def test_qft30_gaussian_three_paths():
    """QFT30 as configured: a seeded i.i.d. complex Gaussian input (normalised),
    H + controlled phase (the reference's controlled R1) + the reversal
    swaps, parts of at most 10 qubits. Per-gate FP64, the part path (parts
    of <= 10 qubits, the reference's gather_and_execute_on), fused FP64 and
    fused FP32-split MFMA (parts of <= 7 qubits, the fused kernels' width)
    from the same input. Bars: the FP64 paths equal per-gate to 1e-12
    relative to the largest amplitude; FP32-split within infidelity 1e-6
    and max error 1e-6 relative."""
    n = 30
    gates = qft(n)
    x = gaussian_state(n, seed=30)
    parts10 = closure_parts(gates, 10)
    parts7 = closure_parts(gates, 7)
    check_parts(gates, parts10, 10)
    check_parts(gates, parts7, 7)
    with capi.State(n) as a, capi.State(n) as b:
        a.write(x)
        a.apply_gates(gates)
        scale = a.max_abs()
        assert abs(a.norm2() - 1) <= TOL
        b.write(x)
        for p in parts10:
            b.apply_part(p, 0)
        assert a.max_abs_diff(b) <= TOL * scale
        b.write(x)
        for p in parts7:
            b.apply_fused(p, "f64")
        assert a.max_abs_diff(b) <= TOL * scale
        assert abs(b.norm2() - 1) <= TOL
        b.write(x)
        for p in parts7:
            b.apply_fused(p, "f32split")
        infid = fused_infidelity(a, b)
        rel = a.max_abs_diff(b) / scale
        print(f"QFT30 FP32-split: infidelity {infid:.3e}, max error {rel:.3e} relative", flush=True)
        assert infid <= 1e-6
        assert rel <= 1e-6


def test_qft30_bf16split_path():
    """The same QFT30 circuit on a dense product-state input through the
    BF16-split fused path (parts of <= 7 qubits) against per-gate FP64:
    infidelity within 1e-6 after renormalisation, unit norm."""
    n = 30
    gates = qft(n)
    prep = [Gate("U3", (q,), (), (0.3 + 0.1 * q, 0.7 * q, 1.1 - 0.05 * q)) for q in range(n)]
    parts = closure_parts(gates, 7)
    with capi.State(n) as a, capi.State(n) as b:
        a.apply_gates(prep + gates)
        b.apply_gates(prep)
        for p in parts:
            b.apply_fused(p, "bf16split")
        b.renormalize()
        assert abs(b.norm2() - 1) <= TOL
        assert fused_infidelity(a, b) <= 1e-6


def test_rl33_depth2_part_and_fused_agree():
    """RL33 (complex128, 137 GB per state) at depth 2: the part path (parts
    of 12 including the 5 low qubits, on the tile kernel) and the fused FP64
    path equal
    the per-gate path to 1e-12 relative to the largest amplitude on two
    full-size states (amplitudes are ~1e-5 here: an absolute bar of 1e-12
    would pass FP32 arithmetic)."""
    n = 33
    gates = random_layered(n, 2, seed=0)
    with capi.State(n) as a, capi.State(n) as b:
        a.apply_gates(gates)
        scale = a.max_abs()
        assert scale > 0
        for p in closure_parts(gates, 12, always=range(5)):
            b.apply_part(p, 0)
        assert a.max_abs_diff(b) <= TOL * scale
        b.reset()
        for p in closure_parts(gates, 6):
            b.apply_fused(p, "f64")
        assert a.max_abs_diff(b) <= TOL * scale


def test_rl34_complex64_depth4():
    """RL34 in complex64 (137 GB per state): BF16-split fused path against
    the complex64 per-gate path, infidelity within 1e-6, unit norm."""
    n = 34
    gates = random_layered(n, 4, seed=0)
    with capi.State(n, "complex64") as a, capi.State(n, "complex64") as b:
        a.apply_gates(gates)
        for p in closure_parts(gates, 7):
            b.apply_fused(p, "bf16split")
        b.renormalize()
        assert abs(b.norm2() - 1) <= 1e-6
        assert fused_infidelity(a, b) <= 1e-6


def test_qaoa_512_batch_closed_form():
    """The 512-instance batch (n ~ U{12..22}, 3-regular graphs, p ~ U{1..4})
    in one hipGraph: every p = 1 instance matches the closed-form <C>."""
    from hisvsim_amd.qaoa import QAOABatch, expectation_p1, make_batch
    insts = make_batch(512, seed=0)
    with QAOABatch(insts) as b:
        got = b.run()
    ones = [i for i, x in enumerate(insts) if x.p == 1]
    assert len(ones) > 50
    for i in ones:
        x = insts[i]
        want = expectation_p1(x.n, x.edges, x.gammas[0], x.betas[0])
        assert abs(got[i] - want) <= 1e-5 * max(1.0, abs(want))


def test_three_piece_qft_error_profile():
    """QFT20 (with the reversal swaps) on a seeded Gaussian input through the
    BF16 chain (k = 13), against per-gate FP64. The three-piece split is
    FP32-class in norm (infidelity <= 2e-13; measured 4.4e-14) and on every
    amplitude (max error <= 1e-6 relative to the largest amplitude; measured
    2.3e-7). Without the kernel's column signs the matrix cores' FP32
    accumulation bias concentrated the error on a few amplitudes (7.8e-14,
    2.5e-6). The two-piece split: infidelity <= 1e-8 (measured 2.1e-10),
    max error <= 1e-4 (1.6e-5)."""
    n = 20
    gates = qft(n)
    x = gaussian_state(n, seed=n)
    parts, lows = chain_parts(gates, 13, beam=(16, 8))
    with capi.State(n) as a, capi.State(n) as b:
        a.write(x)
        a.apply_gates(gates)
        scale = a.max_abs()
        for mode, infid_bar, err_bar in (("bf16x3", 2e-13, 1e-6), ("bf16split", 1e-8, 1e-4)):
            b.write(x)
            b.apply_fused_chain(parts, lows, mode, renormalize=True)
            infid = fused_infidelity(a, b)
            err = a.max_abs_diff(b) / scale
            print(f"QFT20 chain {mode}: infidelity {infid:.2e}, max error {err:.2e} relative", flush=True)
            assert infid <= infid_bar
            assert err <= err_bar


def test_three_piece_error_has_no_coherent_component():
    """The first chain part of QFT20 (no reversal swaps) on two independent
    Gaussian inputs through the three-piece split, against per-gate FP64;
    the error's mean over the 2^7 values of the top seven qubits. Without
    the kernel's column signs the matrix cores' one-directional FP32
    accumulation bias repeated across those values: the two inputs' means
    correlated at 0.97 and were 5.3x the size independent errors leave
    (rms / sqrt(128)). With them: correlation <= 0.3 (measured 0.08;
    independent vectors of 8192 entries: ~0.01) and at most 2x that size
    (measured 1.00x)."""
    n = 20
    gates = [g for g in qft(n) if g.name.upper() != "SWAP"]
    parts, _ = chain_parts(gates, 13, beam=(16, 8))
    means, ratios = [], []
    with capi.State(n) as a, capi.State(n) as s:
        for seed in (20, 99):
            x = gaussian_state(n, seed=seed)
            a.write(x)
            a.apply_gates(parts[0])
            s.write(x)
            s.apply_fused(parts[0], "bf16x3")
            e = (s.read() - a.read()).reshape(2 ** 7, 2 ** 13)
            m = e.mean(axis=0)
            rms = np.sqrt(np.mean(np.abs(e) ** 2))
            means.append(m)
            ratios.append(np.sqrt(np.mean(np.abs(m) ** 2)) / (rms / np.sqrt(2 ** 7)))
    corr = abs(np.vdot(means[0], means[1])) / np.sqrt(np.vdot(means[0], means[0]).real *
                                                      np.vdot(means[1], means[1]).real)
    print(f"QFT20 part 0, three pieces: mean-error correlation {corr:.3f}, "
          f"size {ratios[0]:.2f}x / {ratios[1]:.2f}x of independent errors", flush=True)
    assert corr <= 0.3
    assert max(ratios) <= 2.0
