"""ctypes binding of the libhsv C ABI (include/hsv.h).

This is the same boundary a non-Python caller binds; tests and bench.py go
through it. Every call raises ``HsvError`` on a negative status. There is no
fallback: if libhsv.so is missing or no GPU is present, the calls fail.
"""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, Sequence, Tuple

import numpy as np

from .circuit import Gate

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhsv.so")

HSV_C128, HSV_C64 = 0, 1
HSV_FUSED_F64, HSV_FUSED_F32SPLIT, HSV_FUSED_BF16SPLIT, HSV_FUSED_BF16X3 = 0, 1, 2, 3
CHAIN_RENORMALIZE = 0x100   # HSV_CHAIN_RENORMALIZE
FUSED_MODES = {"f64": HSV_FUSED_F64, "f32split": HSV_FUSED_F32SPLIT, "bf16split": HSV_FUSED_BF16SPLIT,
               "bf16x3": HSV_FUSED_BF16X3}

# SvSim::GateType numbering (basic_gates.hpp:12-45), as in hsv.h.
GATE_TYPES = {
    "I": 3, "X": 4, "Y": 5, "Z": 6, "H": 7, "S": 8, "T": 9, "RX": 10, "RY": 11, "RZ": 12,
    "RI": 13, "R1": 14, "RXFRAC": 15, "RYFRAC": 16, "RZFRAC": 17, "RIFRAC": 18, "R1FRAC": 19,
    "SWAP": 20, "C1": 21, "C2": 22, "U1": 23, "U2": 24, "U3": 25, "RXX": 26, "RYY": 27,
    "RZZ": 28, "TDG": 29, "SDG": 30, "MATRIX": 64,
}
FRAC = {"RXFRAC", "RYFRAC", "RZFRAC", "RIFRAC", "R1FRAC"}

STATUS = {-1: "HSV_ERR_ARG", -2: "HSV_ERR_GATE", -3: "HSV_ERR_ALLOC", -4: "HSV_ERR_DEVICE",
          -5: "HSV_ERR_LIMIT", -6: "HSV_ERR_UNITARY"}


class HsvError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class hsv_gate(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("num_controls", ctypes.c_int32), ("num_targets", ctypes.c_int32),
                ("controls", ctypes.c_int32 * 5), ("targets", ctypes.c_int32 * 5), ("params", ctypes.c_double * 32)]


def to_hsv_gate(g) -> hsv_gate:
    d = hsv_gate()
    name = g.name.upper()
    if name not in GATE_TYPES:
        raise ValueError(f"unknown gate {g.name}")
    d.type = GATE_TYPES[name]
    d.num_controls, d.num_targets = len(g.controls), len(g.targets)
    if d.num_controls > 5 or d.num_targets > 5:
        raise ValueError("at most 5 controls and 5 targets")
    for i, c in enumerate(g.controls):
        d.controls[i] = int(c)
    for i, t in enumerate(g.targets):
        d.targets[i] = int(t)
    if name in FRAC:
        d.params[0], d.params[2] = float(g.params[0]), float(g.params[1])
    else:
        for i, p in enumerate(g.params):
            p = complex(p)
            d.params[2 * i], d.params[2 * i + 1] = p.real, p.imag
    return d


def gate_array(gates: Sequence) -> ctypes.Array:
    arr = (hsv_gate * max(1, len(gates)))()
    for i, g in enumerate(gates):
        arr[i] = to_hsv_gate(g)
    return arr


_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_G = ctypes.POINTER(hsv_gate)

# name -> (restype, argtypes); mirrors include/hsv.h
SIGNATURES = {
    "hsv_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_P)]),
    "hsv_destroy": (ctypes.c_int, [_P]),
    "hsv_num_qubits": (ctypes.c_int, [_P]),
    "hsv_dtype": (ctypes.c_int, [_P]),
    "hsv_apply_gate": (ctypes.c_int, [_P, _G]),
    "hsv_apply_part": (ctypes.c_int, [_P, _G, ctypes.c_int, ctypes.c_int]),
    "hsv_apply_fused": (ctypes.c_int, [_P, _G, ctypes.c_int, ctypes.c_int]),
    "hsv_apply_fused_chain": (ctypes.c_int, [_P, _G, ctypes.c_int, _I, ctypes.c_int, _I, ctypes.c_int]),
    "hsv_apply_fused_chain_noalloc": (ctypes.c_int, [_P, _G, ctypes.c_int, _I, ctypes.c_int, _I, ctypes.c_int]),
    "hsv_chain_parts": (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint64), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), _I]),
    "hsv_part_steps": (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint64), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_int32), _I]),
    "hsv_permute_qubits": (ctypes.c_int, [_P, _I]),
    "hsv_read": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint64, _D]),
    "hsv_write": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint64, _D]),
    "hsv_reset": (ctypes.c_int, [_P]),
    "hsv_norm2": (ctypes.c_int, [_P, _D]),
    "hsv_inner": (ctypes.c_int, [_P, _P, _D]),
    "hsv_max_abs_diff": (ctypes.c_int, [_P, _P, _D]),
    "hsv_probability": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _D]),
    "hsv_measure": (ctypes.c_int, [_P, _I, ctypes.c_int, _D, ctypes.c_int, ctypes.POINTER(ctypes.c_int32)]),
    "hsv_project": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint64, _D]),
    "hsv_renormalize": (ctypes.c_int, [_P, _D]),
    "hsv_expect_pauli": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint64, _D]),
    "hsv_gate_matrix": (ctypes.c_int, [_G, _D, _I]),
    "hsv_fused_unitary": (ctypes.c_int, [ctypes.c_int, _G, ctypes.c_int, _I, _I, _D, _D]),
    "hsv_synchronize": (ctypes.c_int, [_P]),
    "hsv_tune": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int]),
    "hsv_mlp_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(_P)]),
    "hsv_mlp_destroy": (ctypes.c_int, [_P]),
    "hsv_mlp_num_params": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "hsv_mlp_get_params": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_float)]),
    "hsv_mlp_set_params": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_float)]),
    "hsv_mlp_forward": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_float)]),
    "hsv_mlp_gradient": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                        ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "hsv_mlp_train": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]),
    "hsv_qaoa_create": (ctypes.c_int, [ctypes.c_int, _I, _I, _I, _I, ctypes.POINTER(_P)]),
    "hsv_qaoa_destroy": (ctypes.c_int, [_P]),
    "hsv_qaoa_set_angles": (ctypes.c_int, [_P, _D, _D]),
    "hsv_qaoa_run": (ctypes.c_int, [_P, _D]),
    "hsv_qaoa_tune": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int]),
    "hsv_qaoa_last_ms": (ctypes.c_int, [_P, _D]),
    "hsv_qaoa_read_state": (ctypes.c_int, [_P, ctypes.c_int, _D]),
    "hsv_qaoa_total_amplitudes": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64)]),
    "hsv_batch_create": (ctypes.c_int, [ctypes.c_int, _I, _I, _G, ctypes.c_int, ctypes.POINTER(_P)]),
    "hsv_batch_destroy": (ctypes.c_int, [_P]),
    "hsv_batch_set_gates": (ctypes.c_int, [_P, _G]),
    "hsv_batch_run": (ctypes.c_int, [_P]),
    "hsv_batch_expect_z": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64), _D]),
    "hsv_batch_read": (ctypes.c_int, [_P, ctypes.c_int, _D]),
    "hsv_batch_tune": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int]),
    "hsv_batch_info": (ctypes.c_int, [_P, _D, _I, ctypes.POINTER(ctypes.c_uint64)]),
    "hsv_set_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "hsv_kernel_times": (ctypes.c_int, [_P, _D, ctypes.c_int, _I]),
    "hsv_set_opt_slots": (None, [ctypes.c_int]),
    "hsv_get_opt_slots": (ctypes.c_int, []),
    "hsv_get_counters": (None, [ctypes.POINTER(ctypes.c_longlong)] * 4),
    "hsv_last_error": (ctypes.c_char_p, []),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Loads libhsv.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} not built: run `python -m hisvsim_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise HsvError(rc, lib().hsv_last_error().decode())


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_D)


def pauli_masks(spec: str):
    """``"X0 Z3 Y5"`` -> (x_mask, z_mask): X sets the x bit, Z the z bit, Y
    both; ``I`` factors and an empty string are the identity."""
    xm = zm = 0
    for tok in spec.split():
        p, q = tok[0].upper(), int(tok[1:])
        if p not in "IXYZ" or q < 0:
            raise ValueError(f"bad Pauli factor {tok!r}")
        bit = 1 << q
        if (xm | zm) & bit:
            raise ValueError(f"qubit {q} appears twice in {spec!r}")
        if p in "XY":
            xm |= bit
        if p in "ZY":
            zm |= bit
    return xm, zm


def gate_matrix(g) -> np.ndarray:
    """The matrix a gate is lowered to (host only)."""
    out = np.zeros(32)
    dim = ctypes.c_int(0)
    d = to_hsv_gate(g)
    _check(lib().hsv_gate_matrix(ctypes.byref(d), _dptr(out), ctypes.byref(dim)))
    m = (out[0::2] + 1j * out[1::2])
    k = dim.value
    return m[:k * k].reshape(k, k)


def chain_parts(gates: Sequence, k: int, low: int = 5, lookahead: int = 7, candidates: int = 24,
                beam: Tuple[int, int] = (1, 1)):
    """hsv_chain_parts (host only): (parts, lows) as partition.chain_parts;
    beam = (width, branch) of the search over low sets, (1, 1) = greedy.
    Identity gates get an empty mask: the chain leaves them out of a part's
    tile."""
    masks = np.array([0 if g.name.upper() == "I" else sum(1 << q for q in set(g.qubits())) for g in gates],
                     dtype=np.uint64)
    part = np.zeros(max(1, len(gates)), np.int32)
    lows = np.zeros(max(1, len(gates)) * low + low, np.int32)
    num = ctypes.c_int(0)
    _check(lib().hsv_chain_parts(masks.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(gates), k, low,
                                 lookahead, candidates, int(beam[0]), int(beam[1]),
                                 part.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                 lows.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(num)))
    parts = [[] for _ in range(num.value)]
    for g, pi in zip(gates, part[:len(gates)]):
        parts[pi].append(g)
    return parts, [list(map(int, lows[p * low:(p + 1) * low])) for p in range(num.value)]


def part_steps(gates: Sequence, k: int = 5, width: int = 16):
    """hsv_part_steps (host only): the gates of a wide BF16-split part in
    the fused steps the kernel applies (identities left out)."""
    gates = [g for g in gates if g.name.upper() != "I"]
    masks = np.array([sum(1 << q for q in set(g.qubits())) for g in gates], dtype=np.uint64)
    grp = np.zeros(max(1, len(gates)), np.int32)
    num = ctypes.c_int(0)
    _check(lib().hsv_part_steps(masks.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(gates), k, width,
                                grp.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(num)))
    steps = [[] for _ in range(num.value)]
    for g, s in zip(gates, grp[:len(gates)]):
        steps[s].append(g)
    return steps


def fused_unitary(n: int, gates: Sequence):
    """(qubits, U, unitarity_error) of the fused part (host only)."""
    arr = gate_array(gates)
    q = (ctypes.c_int * 8)()
    k = ctypes.c_int(0)
    u = np.zeros(2 * 4 ** 7)
    err = ctypes.c_double(0)
    _check(lib().hsv_fused_unitary(n, arr, len(gates), q, ctypes.byref(k), _dptr(u), ctypes.byref(err)))
    dim = 2 ** k.value
    U = (u[0::2] + 1j * u[1::2])[:dim * dim].reshape(dim, dim)
    return tuple(q[: k.value]), U, err.value


class State:
    """An n-qubit state in HBM (one libhsv handle)."""

    def __init__(self, n: int, dtype: str = "complex128"):
        self.n = n
        self.dtype = dtype
        self._h = ctypes.c_void_p()
        dt = {"complex128": HSV_C128, "complex64": HSV_C64}[dtype]
        _check(lib().hsv_create(n, dt, ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h:
            lib().hsv_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    # -- gates ---------------------------------------------------------------
    def apply_gate(self, g) -> None:
        d = to_hsv_gate(g)
        _check(lib().hsv_apply_gate(self._h, ctypes.byref(d)))

    def apply_gates(self, gates: Iterable) -> None:
        for g in gates:
            self.apply_gate(g)

    def apply_part(self, gates: Sequence, opt_slots: int = 1) -> None:
        arr = gate_array(gates)
        _check(lib().hsv_apply_part(self._h, arr, len(gates), int(opt_slots)))

    def apply_fused(self, gates: Sequence, mode: str = "f64") -> None:
        m = FUSED_MODES[mode]
        arr = gate_array(gates)
        _check(lib().hsv_apply_fused(self._h, arr, len(gates), m))

    def apply_fused_chain(self, parts: Sequence[Sequence], lows: Sequence[Sequence[int]],
                          mode: str = "bf16split", renormalize: bool = False, noalloc: bool = False) -> None:
        """Parts in sequence with a moving qubit layout (hsv_apply_fused_chain):
        lows[p] are the 5 qubits in slots 0-4 while part p runs (lows[0] =
        qubits 0-4); each must be a qubit of the previous part or its lows.
        renormalize: end with the state scaled to unit norm
        (HSV_CHAIN_RENORMALIZE). noalloc: through
        hsv_apply_fused_chain_noalloc (no device allocation; parts of at
        most 28 fused steps)."""
        if len(lows) != len(parts):
            raise ValueError("one lows entry per part")
        flat = [g for p in parts for g in p]
        arr = gate_array(flat)
        sizes = (ctypes.c_int * max(1, len(parts)))(*[len(p) for p in parts])
        lw = [int(q) for low in lows for q in low]
        if any(len(low) != 5 for low in lows):
            raise ValueError("lows: 5 qubits per part")
        larr = (ctypes.c_int * max(1, len(lw)))(*lw)
        fn = lib().hsv_apply_fused_chain_noalloc if noalloc else lib().hsv_apply_fused_chain
        _check(fn(self._h, arr, len(flat), sizes, len(parts), larr,
                  FUSED_MODES[mode] | (CHAIN_RENORMALIZE if renormalize else 0)))

    def permute_qubits(self, new_slot: Sequence[int]) -> None:
        arr = (ctypes.c_int * self.n)(*[int(x) for x in new_slot])
        _check(lib().hsv_permute_qubits(self._h, arr))

    # -- data ----------------------------------------------------------------
    def read(self, offset: int = 0, count: int | None = None) -> np.ndarray:
        count = (2 ** self.n - offset) if count is None else count
        out = np.empty(count, dtype=np.complex128)
        _check(lib().hsv_read(self._h, offset, count, out.ctypes.data_as(_D)))
        return out

    def write(self, amps: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(amps, dtype=np.complex128)
        _check(lib().hsv_write(self._h, offset, a.size, a.ctypes.data_as(_D)))

    def reset(self) -> None:
        _check(lib().hsv_reset(self._h))

    def norm2(self) -> float:
        v = ctypes.c_double()
        _check(lib().hsv_norm2(self._h, ctypes.byref(v)))
        return v.value

    def inner(self, other: "State") -> complex:
        v = (ctypes.c_double * 2)()
        _check(lib().hsv_inner(self._h, other._h, v))
        return complex(v[0], v[1])

    def max_abs_diff(self, other: "State") -> float:
        v = ctypes.c_double()
        _check(lib().hsv_max_abs_diff(self._h, other._h, ctypes.byref(v)))
        return v.value

    def max_abs(self) -> float:
        """max |a_i| (FP64 reduction on the device)."""
        v = ctypes.c_double()
        _check(lib().hsv_max_abs_diff(self._h, None, ctypes.byref(v)))
        return v.value

    def rel_diff(self, other: "State") -> float:
        """max |a_i - b_i| / max |a_i|: the error bar that scales with the
        amplitudes (at 30+ qubits they are ~1e-5, where an absolute 1e-12
        could not tell FP32 from FP64 arithmetic)."""
        return self.max_abs_diff(other) / self.max_abs()

    def probability(self, qubit: int, value: int = 0) -> float:
        v = ctypes.c_double()
        _check(lib().hsv_probability(self._h, qubit, value, ctypes.byref(v)))
        return v.value

    def measure(self, qubits: Sequence[int], uniforms: np.ndarray) -> np.ndarray:
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        out = np.zeros((len(u), len(qubits)), dtype=np.int32)
        q = (ctypes.c_int * max(1, len(qubits)))(*qubits)
        _check(lib().hsv_measure(self._h, q, len(qubits), u.ctypes.data_as(_D), len(u),
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return out

    def renormalize(self) -> float:
        """Scales the state to unit norm; returns the norm^2 before."""
        v = ctypes.c_double()
        _check(lib().hsv_renormalize(self._h, ctypes.byref(v)))
        return v.value

    def project(self, qubits: Sequence[int], values: Sequence[int] = ()) -> float:
        """Collapses the qubits to the values (0 where not given, as the
        reference's default), renormalised; returns the outcome's
        probability."""
        values = list(values) + [0] * (len(qubits) - len(values))
        mask = val = 0
        for q, v in zip(qubits, values):
            if v not in (0, 1) or (mask >> q) & 1:
                raise ValueError("values must be 0/1 and qubits unique")
            mask |= 1 << q
            val |= v << q
        p = ctypes.c_double()
        _check(lib().hsv_project(self._h, mask, val, ctypes.byref(p)))
        return p.value

    def expectation(self, pauli) -> float:
        """<psi|P|psi> for a Pauli string given as ``"X0 Z3 Y5"`` (factor
        letter + qubit, identity elsewhere) or as an (x_mask, z_mask) pair."""
        xm, zm = pauli_masks(pauli) if isinstance(pauli, str) else pauli
        v = (ctypes.c_double * 2)()
        _check(lib().hsv_expect_pauli(self._h, int(xm), int(zm), v))
        return v[0]

    def tune(self, key: str, value: int) -> None:
        _check(lib().hsv_tune(self._h, key.encode(), int(value)))

    # -- sync / timing ----------------------------------------------------------
    def synchronize(self) -> None:
        _check(lib().hsv_synchronize(self._h))

    def set_timing(self, on: bool) -> None:
        _check(lib().hsv_set_timing(self._h, int(on)))

    def kernel_times(self, capacity: int = 65536) -> np.ndarray:
        """Durations (ms) of the timed calls since timing was switched on or
        last collected, in call order."""
        out = np.zeros(capacity)
        cnt = ctypes.c_int(0)
        _check(lib().hsv_kernel_times(self._h, out.ctypes.data_as(_D), capacity, ctypes.byref(cnt)))
        return out[:min(cnt.value, capacity)]


def counters() -> dict:
    v = [ctypes.c_longlong() for _ in range(4)]
    lib().hsv_get_counters(*[ctypes.byref(x) for x in v])
    return dict(zip(("apply_us", "gates", "parts", "part_us"), (x.value for x in v)))
