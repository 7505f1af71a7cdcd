"""CPU model of the three-piece BF16 split's arithmetic (TEST INFRASTRUCTURE).

Models what ``k_part_bf`` computes with three pieces per value, so that the
tests can study its rounding on the CPU: the FP32 state of a tile in LDS,
each 5-qubit step as 24 BF16 matrix products accumulated in FP32, and the
column signs the kernel applies around every step. Only tests import it.

- ``split_values``: the device split of FP32 values (split_pair3 in
  fused.hip): hi = BF16(a), mid = BF16(a - hi), lo = BF16(a - hi - mid),
  BF16 conversions rounding to nearest even, the remainders in FP32;
- ``split_matrix``: the host split of the FP64 step unitary (part_apply_bf):
  each piece the BF16 (via FP32) of what FP64 leaves after the previous ones;
- ``step``: Re = Ur Xr - Ui Xi, Im = Ur Xi + Ui Xr, each real product as the
  six piece products smallest first (lo*hi, hi*lo, mid*mid, hi*mid, mid*hi,
  hi*hi) into an FP32 accumulator; one matrix product adds its 32 exact
  BF16 x BF16 products (summed in FP64 here) to the accumulator and rounds
  once, to nearest or toward minus infinity (``rounding``); the final
  Re / Im subtraction and addition round to nearest;
- ``column_hash`` / ``column_sign``: the kernel's sign of a column of a tile
  in one step (fused.hip); ``chain`` applies them around every step when
  ``signs`` is set, over this model's enumeration of tiles (the qubits
  outside the tile, ascending) and columns (the tile's qubits outside the
  step, ascending).

``chain`` runs a partitioned circuit the way ``hsv_apply_fused_chain``
does: per part the state rounds to FP32 (the load into LDS), then the
part's fused steps (``capi.part_steps``, ``capi.fused_unitary``).
"""
from __future__ import annotations

from typing import Sequence

import numpy as np

M32 = np.uint64(0xFFFFFFFF)


def bf16(a: np.ndarray) -> np.ndarray:
    """BF16 of FP32 values, round to nearest even, as FP32 values."""
    b = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = (b + (((b >> np.uint64(16)) & np.uint64(1)) + np.uint64(0x7FFF))) & np.uint64(0xFFFF0000)
    return b.astype(np.uint32).view(np.float32)


def split_values(v: np.ndarray) -> list:
    """Three BF16 pieces of FP32 values; remainders in FP32."""
    out, r = [], np.asarray(v, dtype=np.float32)
    for _ in range(3):
        p = bf16(r)
        out.append(p)
        r = (r - p).astype(np.float32)
    return out


def split_matrix(u: np.ndarray) -> list:
    """Three BF16 pieces of FP64 values; remainders in FP64."""
    out, r = [], np.asarray(u, dtype=np.float64)
    for _ in range(3):
        p = bf16(r.astype(np.float32))
        out.append(p)
        r = r - p.astype(np.float64)
    return out


def _round32(v: np.ndarray, rounding: str) -> np.ndarray:
    f = v.astype(np.float32)
    if rounding == "down":
        f = np.where(f.astype(np.float64) > v, np.nextafter(f, np.float32(-np.inf)), f)
    elif rounding != "nearest":
        raise ValueError(rounding)
    return f


# (U piece, X piece) in accumulation order: smallest products first
ORDER = ((0, 2), (2, 0), (1, 1), (0, 1), (1, 0), (0, 0))


def step(x: np.ndarray, u: np.ndarray, rounding: str = "nearest") -> np.ndarray:
    """Y = X U^T for rows x (complex, FP32 components) and a 32 x 32 (or
    smaller) unitary u, in the kernel's arithmetic."""
    xr, xi = split_values(x.real), split_values(x.imag)
    ur, ui = split_matrix(u.real), split_matrix(u.imag)

    def product(us, xs):
        acc = np.zeros((x.shape[0], u.shape[0]), np.float32)
        for pu, px in ORDER:
            acc = _round32(acc.astype(np.float64) + xs[px].astype(np.float64) @ us[pu].astype(np.float64).T,
                           rounding)
        return acc

    re = (product(ur, xr) - product(ui, xi)).astype(np.float32)
    im = (product(ur, xi) + product(ui, xr)).astype(np.float32)
    return re.astype(np.float64) + 1j * im.astype(np.float64)


def column_hash(key: np.ndarray) -> np.ndarray:
    """murmur3's 32-bit finaliser of (tile << 8 | column)."""
    k = key.astype(np.uint64) & M32
    k ^= k >> np.uint64(16)
    k = (k * np.uint64(0x85EBCA6B)) & M32
    k ^= k >> np.uint64(13)
    k = (k * np.uint64(0xC2B2AE35)) & M32
    k ^= k >> np.uint64(16)
    return k


def column_sign(h: np.ndarray, s: int) -> np.ndarray:
    """+1 / -1: the top bit of (hash ^ step * 0x85EBCA77) * 0x9E3779B1."""
    salt = np.uint64((s * 0x85EBCA77) & 0xFFFFFFFF)
    k = ((h ^ salt) * np.uint64(0x9E3779B1)) & M32
    return np.where(k >> np.uint64(31), -1.0, 1.0)


def _pext(idx: np.ndarray, bits: Sequence[int]) -> np.ndarray:
    out = np.zeros_like(idx)
    for j, b in enumerate(bits):
        out |= ((idx >> np.uint64(b)) & np.uint64(1)) << np.uint64(j)
    return out


def apply_rows(state: np.ndarray, n: int, qubits: Sequence[int], fn) -> np.ndarray:
    """Applies fn(X) -> Y to the rows X[m, j] = state[index with the qubits'
    bits = j (qubits[0] lowest), the others = m]."""
    k = len(qubits)
    t = state.reshape([2] * n)
    axes = [n - 1 - q for q in reversed(qubits)]
    t = np.moveaxis(t, axes, list(range(n - k, n)))
    shape = t.shape
    t = fn(t.reshape(-1, 2 ** k)).reshape(shape)
    return np.ascontiguousarray(np.moveaxis(t, list(range(n - k, n)), axes)).reshape(-1)


def chain(x: np.ndarray, n: int, parts, lows, rounding: str = "nearest", signs: bool = False,
          tile_qubits: int = 13) -> np.ndarray:
    """The partitioned circuit on the modelled three-piece arithmetic."""
    from hisvsim_amd import capi
    idx = np.arange(2 ** n, dtype=np.uint64)
    got = np.asarray(x, dtype=np.complex128)
    for part, low in zip(parts, lows):
        tq = sorted(set(q for g in part for q in g.qubits()) | set(low))
        tq = sorted(tq + [q for q in range(n) if q not in tq][: max(0, tile_qubits - len(tq))])
        tile = _pext(idx, [q for q in range(n) if q not in tq])
        got = got.astype(np.complex64).astype(np.complex128)
        for s, gates in enumerate(capi.part_steps(part)):
            qs, u, _ = capi.fused_unitary(n, gates)
            if signs:
                col = _pext(idx, [q for q in tq if q not in qs])
                sg = column_sign(column_hash((tile << np.uint64(8)) | col), s)
                got = got * sg
            got = apply_rows(got, n, list(qs), lambda rows: step(rows, u, rounding))
            if signs:
                got = got * sg
    return got


def exact(x: np.ndarray, n: int, parts) -> np.ndarray:
    """The same fused steps in FP64."""
    from hisvsim_amd import capi
    ref = np.asarray(x, dtype=np.complex128)
    for part in parts:
        for gates in capi.part_steps(part):
            qs, u, _ = capi.fused_unitary(n, gates)
            ref = apply_rows(ref, n, list(qs), lambda rows: rows @ u.T)
    return ref
