"""The three-piece BF16 split's rounding, on a CPU model of the kernel's
arithmetic (tests/bf16x3_model.py): the column signs k_part_bf applies
around every three-piece step are exact (bit-identical results under
round-to-nearest accumulation), and they take apart the error
concentration that an FP32 accumulation rounding in one direction produces
on a QFT."""
import numpy as np

import bf16x3_model as M
from hisvsim_amd.circuit import gaussian_state, qft
from hisvsim_amd.partition import chain_parts


def test_model_pieces_and_hash():
    """The pieces add up to the value (FP32 values: exactly; FP64 values: to
    2^-24 of it), a step is FP32-class against the FP64 product, and
    column_hash is murmur3's 32-bit finaliser (published values: 0 -> 0,
    1 -> 0x514E28B7)."""
    rng = np.random.default_rng(5)
    v = rng.standard_normal(4096).astype(np.float32)
    hi, mid, lo = M.split_values(v)
    assert np.array_equal(hi.astype(np.float64) + mid + lo, v.astype(np.float64))
    u = rng.standard_normal((32, 32))
    assert np.abs(sum(p.astype(np.float64) for p in M.split_matrix(u)) - u).max() <= 2 ** -24 * np.abs(u).max()
    q, _ = np.linalg.qr(rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32)))
    x = (rng.standard_normal((256, 32)) + 1j * rng.standard_normal((256, 32))).astype(np.complex64)
    want = x.astype(np.complex128) @ q.T
    got = M.step(x.astype(np.complex128), q)
    assert np.abs(got - want).max() <= 2e-7 * np.abs(want).max()   # measured 5.1e-8
    # each of the six products matters: without any one of them the error
    # is at least 2.6e-6 (a smallest product) and up to 1 (hi * hi)
    keep = M.ORDER
    try:
        for i in range(len(keep)):
            M.ORDER = keep[:i] + keep[i + 1:]
            assert np.abs(M.step(x.astype(np.complex128), q) - want).max() > 1e-6 * np.abs(want).max()
    finally:
        M.ORDER = keep
    h = M.column_hash(np.array([0, 1], dtype=np.uint64))
    assert [int(v) for v in h] == [0, 0x514E28B7]


def test_row_signs_flip_a_directed_bias_only():
    """s * step(s * X) equals step(X) bit for bit when the accumulation rounds
    to nearest (rounding commutes with negation), and not when it rounds
    toward minus infinity."""
    rng = np.random.default_rng(7)
    q, _ = np.linalg.qr(rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32)))
    x = (rng.standard_normal((512, 32)) + 1j * rng.standard_normal((512, 32))).astype(np.complex64)
    x = x.astype(np.complex128)
    s = np.where(rng.random(512) < 0.5, -1.0, 1.0)[:, None]
    assert np.array_equal(s * M.step(s * x, q, "nearest"), M.step(x, q, "nearest"))
    assert not np.array_equal(s * M.step(s * x, q, "down"), M.step(x, q, "down"))


def test_column_signs_spread_a_directed_accumulation_bias_on_qft():
    """QFT16 on a seeded Gaussian input through the three-piece chain (k =
    13), against the same fused steps in FP64. Rounding to nearest: largest
    error 4.6x the median, and the column signs change no bit. Rounding every
    accumulation toward minus infinity: the error concentrates (largest 26x
    the median, 1.9e-6 relative to the largest amplitude); with the kernel's
    column signs it is spread again (5.1x, 2.8e-7)."""
    n = 16
    gates = qft(n)
    parts, lows = chain_parts(gates, 13, beam=(16, 8))
    x = gaussian_state(n, seed=n)
    ref = M.exact(x, n, parts)
    scale = np.abs(ref).max()

    def profile(got):
        e = np.abs(got - ref) / scale
        return e.max(), e.max() / np.median(e)

    near = M.chain(x, n, parts, lows, "nearest", signs=False)
    assert np.array_equal(M.chain(x, n, parts, lows, "nearest", signs=True), near)
    down_max, down_ratio = profile(M.chain(x, n, parts, lows, "down", signs=False))
    flip_max, flip_ratio = profile(M.chain(x, n, parts, lows, "down", signs=True))
    near_max, near_ratio = profile(near)
    print(f"QFT16 model: nearest {near_max:.2e} ({near_ratio:.1f}x median), down {down_max:.2e} "
          f"({down_ratio:.1f}x), down with column signs {flip_max:.2e} ({flip_ratio:.1f}x)")
    assert near_ratio <= 8
    assert down_ratio >= 15
    assert flip_ratio <= 8
    assert flip_max <= down_max / 4


def test_model_hash_matches_the_kernel_source():
    """The model's column hash and sign use the constants of column_hash and
    column_sign in fused.hip, so the model studies the kernel's signs."""
    import os
    import re
    src = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            "hisvsim_amd", "csrc", "fused.hip")).read()
    body_hash = src[src.index("uint32_t column_hash("):]
    body_hash = body_hash[:body_hash.index("\n}")]
    body_sign = src[src.index("uint32_t column_sign("):]
    body_sign = body_sign[:body_sign.index("\n}")]
    consts = lambda body: [int(c, 16) for c in re.findall(r"0x([0-9A-Fa-f]+)u", body)]
    assert consts(body_hash) == [0x85EBCA6B, 0xC2B2AE35]
    assert consts(body_sign) == [0x85EBCA77, 0x9E3779B1, 0x80000000]
    assert "<< 8" in body_hash
    # and the model's
    import inspect
    model_consts = lambda fn: [int(c, 16) for c in re.findall(r"0x([0-9A-Fa-f]+)", inspect.getsource(fn))]
    assert set(model_consts(M.column_hash)) == {0x85EBCA6B, 0xC2B2AE35}
    assert set(model_consts(M.column_sign)) - {0xFFFFFFFF} == {0x85EBCA77, 0x9E3779B1}
