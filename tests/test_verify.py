"""Inverse circuits for the driver's accuracy check, on the CPU oracle:
every unitary gate of the reference set (with 0-2 controls) followed by its
inverse gives back the input state to 1e-12, whole random circuits too;
gates without an exact inverse are refused."""
import numpy as np
import pytest

from conftest import rand_state
from gatesets import UNITARY_ONE, UNITARY_TWO, params_for, random_circuit
from hisvsim_amd.circuit import Gate, qft, random_layered
from hisvsim_amd.verify import inverse_circuit, inverse_gate


def unitary_c1(rng):
    q, r = np.linalg.qr(rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2)))
    return tuple(complex(x) for x in (q * (np.diag(r) / np.abs(np.diag(r)))).reshape(-1))


@pytest.mark.parametrize("name", [g for g in UNITARY_ONE + UNITARY_TWO if g != "I"] + ["C1"])
def test_inverse_gate_on_the_oracle(orc, name):
    rng = np.random.default_rng(abs(hash(name)) % 1000)
    n = 5
    for _ in range(6):
        nt = 2 if name in UNITARY_TWO else 1
        nc = int(rng.integers(0, 3))
        qs = [int(q) for q in rng.choice(n, nt + nc, replace=False)]
        params = unitary_c1(rng) if name == "C1" else params_for(name, rng)
        g = Gate(name, tuple(qs[:nt]), tuple(qs[nt:]), params)
        a = rand_state(n, int(rng.integers(1000)))
        b = a.copy()
        orc.apply_gates(b, [g] + inverse_gate(g))
        assert np.abs(b - a).max() <= 1e-12


def test_inverse_circuits_on_the_oracle(orc):
    rng = np.random.default_rng(3)
    for gates, n in ((random_circuit(rng, 7, 120), 7), (random_layered(8, 6, seed=2), 8), (qft(7), 7)):
        a = rand_state(n, 11)
        b = a.copy()
        orc.apply_gates(b, list(gates) + inverse_circuit(gates))
        assert np.abs(b - a).max() <= 1e-12


def test_gates_without_an_exact_inverse_are_refused():
    rng = np.random.default_rng(0)
    with pytest.raises(ValueError):
        inverse_gate(Gate("C2", (0, 1), (), params_for("C2", rng)))   # the reference's C2 matrix is singular
    with pytest.raises(ValueError):
        inverse_gate(Gate("C1", (0,), (), (1, 2, 3, 4)))               # not unitary


def test_inverse_of_every_shipped_circuit(orc):
    """Every QASM fixture, in both lowering modes, has an exact inverse;
    those of up to 12 qubits go back to |0...0> on the oracle."""
    import glob
    import os
    from hisvsim_amd import qasm
    files = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "qasm", "*.qasm")))
    assert files
    for path in files:
        prog = qasm.parse_file(path)
        for mode in ("reference", "textbook"):
            gates = prog.gates(mode)
            inv = inverse_circuit(gates)
            n = prog.num_qubits
            if n <= 12:
                a = np.zeros(2 ** n, complex)
                a[0] = 1
                orc.apply_gates(a, list(gates) + inv)
                assert abs(a[0] - 1) <= 1e-12 and np.abs(a[1:]).max() <= 1e-12, (path, mode)
