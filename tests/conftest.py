import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu on the GPU host)")


@pytest.fixture(scope="session")
def orc():
    from oracle.oracle import Oracle
    return Oracle()


def rand_state(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    return a / np.linalg.norm(a)


def bit_reverse_perm(n: int) -> np.ndarray:
    idx = np.arange(2 ** n)
    rev = np.zeros_like(idx)
    for b in range(n):
        rev |= ((idx >> b) & 1) << (n - 1 - b)
    return rev


def layered_inverse(gates):
    """The inverse of a random layered circuit (U3 and CZ gates): the gates
    in reverse order, U3(theta, phi, lam) undone by U3(-theta, -lam, -phi)
    (in the reference's convention, U3 = exp(-i (phi + lam) / 2) times the
    textbook matrix, this is the exact inverse, phase included), CZ by
    itself."""
    from hisvsim_amd.circuit import Gate
    out = []
    for g in reversed(gates):
        if g.name == "U3":
            th, ph, lam = g.params
            out.append(Gate("U3", g.targets, g.controls, (-th, -lam, -ph)))
        elif g.name == "CZ" or (g.name == "Z" and len(g.controls) == 1):
            out.append(g)
        else:
            raise ValueError(f"no inverse for {g.name}")
    return out
