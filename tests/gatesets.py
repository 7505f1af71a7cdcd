"""Gate-set helpers shared by the tests: every reference gate with random
parameters, random circuits and random acyclic parts."""
import math

import numpy as np

from hisvsim_amd.circuit import Gate

ONE_NOPARAM = ["I", "X", "Y", "Z", "H", "S", "T", "TDG", "SDG"]
ONE_PARAM = ["RX", "RY", "RZ", "RI", "R1", "U1"]
FRAC = ["RXFRAC", "RYFRAC", "RZFRAC", "RIFRAC", "R1FRAC"]
TWO_PARAM = ["RXX", "RYY", "RZZ"]
ALL_ONE = ONE_NOPARAM + ONE_PARAM + ["U2", "U3", "C1"] + FRAC
ALL_TWO = ["SWAP", "C2"] + TWO_PARAM
UNITARY_ONE = [g for g in ALL_ONE if g != "C1"]
UNITARY_TWO = ["SWAP"] + TWO_PARAM


def params_for(name: str, rng: np.random.Generator) -> tuple:
    if name in ONE_PARAM or name in TWO_PARAM:
        return (float(rng.uniform(0, 2 * math.pi)),)
    if name == "U2":
        return tuple(float(x) for x in rng.uniform(0, 2 * math.pi, 2))
    if name == "U3":
        return tuple(float(x) for x in rng.uniform(0, 2 * math.pi, 3))
    if name == "C1":
        return tuple(complex(a, b) for a, b in rng.standard_normal((4, 2)))
    if name == "C2":
        return tuple(complex(a, b) for a, b in rng.standard_normal((16, 2)))
    if name in FRAC:
        return (int(rng.integers(0, 9)), int(rng.integers(1, 7)))
    return ()


def random_gate(rng, n: int, names=None, max_controls: int = 2) -> Gate:
    names = names or (UNITARY_ONE + UNITARY_TWO)
    name = names[int(rng.integers(len(names)))]
    nt = 2 if name in ALL_TWO else 1
    nc = int(rng.integers(0, min(max_controls, n - nt) + 1))
    qs = rng.choice(n, size=nt + nc, replace=False)
    return Gate(name, tuple(int(q) for q in qs[:nt]), tuple(int(q) for q in qs[nt:]), params_for(name, rng))


def random_circuit(rng, n: int, ngates: int, names=None, max_controls: int = 2):
    return [random_gate(rng, n, names, max_controls) for _ in range(ngates)]


def random_parts(rng, n: int, nparts: int, k: int, gates_per_part: int, names=None, max_controls: int = 1):
    """Consecutive parts, each with all its gates inside a random k-qubit set."""
    parts = []
    for _ in range(nparts):
        qs = [int(q) for q in rng.choice(n, size=k, replace=False)]
        part = []
        for _ in range(gates_per_part):
            g = random_gate(rng, k, names, max_controls)
            part.append(Gate(g.name, tuple(qs[t] for t in g.targets), tuple(qs[c] for c in g.controls), g.params))
        parts.append(part)
    return parts
