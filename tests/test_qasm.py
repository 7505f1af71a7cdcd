"""QASM front end: the ordered list, the ``_exec`` dump and the DAG files
against the reference's own files (byte for byte, via SHA-256), the parts
files (natural-order partition reproduced exactly, shipped partitions valid),
parser semantics, lowering, and closed-form circuit results (CPU oracle; GPU
driver runs marked ``gpu``)."""
import hashlib
import json
import math
import os

import numpy as np
import pytest

from hisvsim_amd import qasm
from hisvsim_amd.circuit import Gate
from hisvsim_amd.partition import check_parts, nat_part_ids

QDIR = os.path.join(os.path.dirname(__file__), "golden", "qasm")
FIX = json.load(open(os.path.join(QDIR, "qasm_fixtures.json")))


def prog_of(name):
    return qasm.parse_file(os.path.join(QDIR, name + ".qasm"))


def sha(text: str) -> str:
    return hashlib.sha256(text.encode()).hexdigest()


@pytest.mark.parametrize("dot", sorted(FIX["dot"]))
def test_dag_file_matches_reference(dot):
    """qasm_assembler_standalone.py:3688-3764 output, byte for byte."""
    want = FIX["dot"][dot]
    circuit, part = dot.split(".qasm_PART_")
    part = int(part.split(".")[0])
    segs = [s for s in qasm.build_dags(prog_of(circuit).ops) if s.dag.nodes]
    if want["bytes"] == 0:
        # the shipped file is empty (the reference run did not finish writing it)
        assert len(segs) > part
        return
    assert sha(segs[part].dag.to_dot()) == want["sha256"]


@pytest.mark.parametrize("dump", sorted(FIX["exec"]))
def test_exec_dump_matches_reference(dump):
    """qasm_assembler_standalone.py:410-423 output, byte for byte."""
    got = qasm.exec_dump(prog_of(dump[: -len(".qasm_exec")]))
    assert sha(got) == FIX["exec"][dump]["sha256"] and len(got) == FIX["exec"][dump]["bytes"]


NAT_K = {"bv_n30": 16, "qft_n26": 16, "qft_n30": 18}


@pytest.mark.parametrize("circuit", sorted(NAT_K))
def test_natural_order_parts_reproduce_reference_files(circuit):
    """The shipped ``*_part_nat`` files are the natural-order greedy partition
    (new part when the next gate would exceed k qubits) written in DAG node
    order, input/exit nodes in the part of their neighbouring gate."""
    prog = prog_of(circuit)
    ids = nat_part_ids(prog.gates(), NAT_K[circuit])
    part_of_order = dict(zip(qasm.gate_order_ids(prog), ids))
    seg = qasm.build_dags(prog.ops)[0]
    got = qasm.parts_file_text(seg, part_of_order).splitlines()
    want = open(os.path.join(QDIR, f"{circuit}_part_nat")).read().splitlines()
    if circuit == "bv_n30":
        # this parts file was written from a version of bv_n30.qasm with 29
        # measurements (its exit nodes carry order 101 + 8192, the .qasm and
        # its _exec dump hold 72 operations): compare exit orders modulo that
        exit_order = len(prog.ops) + 8192
        want = [" ".join([str(exit_order), *w.split()[1:]]) if "_exit_" in w else w for w in want]
    assert got == want


@pytest.mark.parametrize("pf", sorted(FIX["parts"]))
def test_shipped_partitions_are_valid(pf):
    """Every shipped partition (natural order, DFS order, dagP 'smart') keeps
    each qubit's gate order when run in ascending part id (the acyclicity
    the part executor relies on)."""
    circuit = pf.rsplit("_part_", 1)[0]
    prog = prog_of(circuit)
    parts = qasm.parts_from_parts_file(prog, os.path.join(QDIR, pf))
    check_parts(prog.gates(), parts)


def test_parser_semantics():
    text = """OPENQASM 2.0;
include "qelib1.inc";
qreg a[2]; qreg b[3];
creg c[3];
gate rot(theta, phi) x, y { rz(theta/2) x; cx x, y; ry(-phi) y; }
h a; cx a[0],b[1]; // trailing comment ; with a semicolon
u3(3*pi/4, 1.5e-3, -sin(pi/2)) b[2];
rot(pi, 2^2) a[1], b[0];
barrier a, b;
cx a, b[2];
measure b -> c;
"""
    p = qasm.parse(text)
    assert list(p.qregs.items()) == [("a", 2), ("b", 3)]
    assert p.qubit_ids() == {"a[0]": 0, "a[1]": 1, "b[0]": 2, "b[1]": 3, "b[2]": 4}
    got = [(o.gate, o.params, o.qubits, o.index) for o in p.ops]
    want = [
        ("h", [], ["a[0]"], 5), ("h", [], ["a[1]"], 5), ("cx", [], ["a[0]", "b[1]"], 5),
        ("u3", [3 * math.pi / 4, 1.5e-3, -1.0], ["b[2]"], 6),
        ("rz", [math.pi / 2], ["a[1]"], 4), ("cx", [], ["a[1]", "b[0]"], 4), ("ry", [-4.0], ["b[0]"], 4),
        ("cx", [], ["a[0]", "b[2]"], 9), ("cx", [], ["a[1]", "b[2]"], 9),
        ("measure", [], ["b[0]", "c[0]"], 10), ("measure", [], ["b[1]", "c[1]"], 10),
        ("measure", [], ["b[2]", "c[2]"], 10),
    ]
    assert got == want


@pytest.mark.parametrize("text", [
    "qreg q[2]; if(c==1) x q[0];",
    "qreg q[2]; foo q[0];",
    "qreg q[2]; x r[0];",
    "qreg q[2]; qreg r[3]; cx q, r;",
    "qreg q[2]; rx(pi/) q[0];",
    "qreg q[2]; rx(__import__) q[0];",
    "qreg q[1]; gate g(t) a { rx(t) a; } g q[0];",
])
def test_parser_rejects(text):
    with pytest.raises(qasm.QasmError):
        qasm.parse(text).gates()


def test_lowering_table():
    """qasm_assembler_standalone.py:1803-2010, with SWAP, CP, SX fixed."""
    p = qasm.parse("""qreg q[5];
cx q[0],q[1]; cz q[2],q[0]; cu1(0.5) q[3],q[4]; cu3(0.1,0.2,0.3) q[1],q[2]; crz(0.7) q[0],q[4];
rzz(0.25) q[1],q[3]; ccx q[0],q[1],q[2]; cswap q[4],q[0],q[1]; swap q[2],q[3]; cp(0.9) q[0],q[1];
u1(0.4) q[0]; u2(0.1,0.2) q[1]; id q[2]; sx q[3];""")
    g = p.gates()
    assert g[:10] == [
        Gate("X", (1,), (0,)), Gate("Z", (0,), (2,)), Gate("U1", (4,), (3,), (0.5,)),
        Gate("U3", (2,), (1,), (0.1, 0.2, 0.3)), Gate("RZ", (4,), (0,), (0.7,)),
        Gate("RZZ", (1, 3), (), (0.25,)), Gate("X", (2,), (0, 1)), Gate("SWAP", (0, 1), (4,)),
        Gate("SWAP", (2, 3)), Gate("R1", (1,), (0,), (0.9,)),
    ]
    assert g[10:13] == [Gate("U1", (0,), (), (0.4,)), Gate("U2", (1,), (), (0.1, 0.2)), Gate("I", (2,))]
    assert g[13].name == "C1" and np.allclose(np.array(g[13].params).reshape(2, 2) @ np.array(
        g[13].params).reshape(2, 2), [[0, 1], [1, 0]])
    t = p.gates("textbook")
    assert t[2] == Gate("R1", (4,), (3,), (0.5,)) and t[10] == Gate("R1", (0,), (), (0.4,))
    th, ph, la = 0.1, 0.2, 0.3
    want = np.array([[math.cos(th / 2), -np.exp(1j * la) * math.sin(th / 2)],
                     [np.exp(1j * ph) * math.sin(th / 2), np.exp(1j * (ph + la)) * math.cos(th / 2)]])
    assert t[3].name == "C1" and np.allclose(np.array(t[3].params).reshape(2, 2), want)


def test_lowering_rejects_reset():
    with pytest.raises(qasm.QasmError):
        qasm.parse("qreg q[1]; reset q[0];").gates()


def bv_mask(prog):
    """Hidden string of a BV circuit: the controls of the CX gates onto the
    ancilla (the last qubit)."""
    n = prog.num_qubits
    return sum(1 << g.controls[0] for g in prog.gates() if g.name == "X" and g.controls and g.targets[0] == n - 1)


def closed_form(name, prog):
    """Expected final state (up to a global phase) for circuits with a closed
    form: BV (data = mask, ancilla |->), QFT of |0> (uniform), X-only
    circuits (a basis state)."""
    n = prog.num_qubits
    if name.startswith("bv") or name == "simple":
        m = bv_mask(prog)
        s = np.zeros(2 ** n, complex)
        s[m], s[m + 2 ** (n - 1)] = 1 / math.sqrt(2), -1 / math.sqrt(2)
        return s
    if name.startswith("qft"):
        return np.full(2 ** n, 2 ** (-n / 2), complex)
    if name.startswith("q_0_25"):
        idx = 0
        for g in prog.gates():
            assert g.name == "X" and not g.controls
            idx ^= 1 << g.targets[0]
        s = np.zeros(2 ** n, complex)
        s[idx] = 1
        return s
    return None


def phase_aligned_diff(got, want):
    k = int(np.argmax(np.abs(want)))
    ph = got[k] / want[k]
    return float(np.abs(got - ph * want).max()), abs(abs(ph) - 1)


@pytest.mark.parametrize("name", ["bv_n14", "bv_n19", "simple", "qft_n15"])
def test_closed_forms_on_oracle(orc, name):
    """The lowered circuits on the CPU oracle reproduce the closed forms, in
    both lowering modes (u1 differs from the phase gate by a global phase
    only when uncontrolled)."""
    prog = prog_of(name)
    want = closed_form(name, prog)
    for mode in ("reference", "textbook"):
        s = np.zeros(2 ** prog.num_qubits, complex)
        s[0] = 1
        orc.apply_gates(s, prog.gates(mode))
        d, dn = phase_aligned_diff(s, want)
        assert d <= 1e-12 and dn <= 1e-12


# ---- GPU: the driver end to end ------------------------------------------------------
def _args(**kw):
    import argparse
    base = dict(app="", parts=False, partfile="", nat=0, greedy=0, slots=False, mode="part", layout="fixed",
                beam="4,2", plan="file", lowering="reference", dtype="auto", dot=False, dump=False, shots=10,
                seed=0, n=0, alloc=0, initn=0, verify="auto")
    base.update(kw)
    return argparse.Namespace(**base)


@pytest.mark.gpu
@pytest.mark.parametrize("name,partfile,extra", [
    ("bv_n19", "bv_n19_part_smart", {}),
    ("bv_n26", "bv_n26_part_smart", {"slots": True}),
    ("bv_n30", "bv_n30_part_nat", {}),
    ("bv_n30", "bv_n30_part_dfs", {}),
    ("bv_n30", "", {"greedy": 6, "mode": "fused64"}),
    ("bv_n26", "", {"greedy": 5, "mode": "fused32"}),
    ("simple", "", {"nat": 8}),
    ("q_0_25_1", "", {"greedy": 10}),
    ("bv_n26", "", {"mode": "fusedbf16", "layout": "chain"}),
    ("bv_n30", "", {"mode": "fusedbf16", "layout": "chain", "nat": 12}),
    ("bv_n26", "", {"plan": "auto"}),
    ("q_0_25_1", "", {"plan": "auto"}),
])
def test_driver_closed_form_circuits(name, partfile, extra):
    """Driver runs of the reference's circuits with their shipped parts files
    (or this package's partitioners): BV gives its hidden string with
    certainty, the X-only circuit its basis state."""
    from hisvsim_amd import driver
    prog = prog_of(name)
    pf = os.path.join(QDIR, partfile) if partfile else ""
    out = driver.run(_args(app=os.path.join(QDIR, name + ".qasm"), parts=True, partfile=pf, **extra))
    # FP64 paths: 1e-12; the split-precision fused paths: their 1e-6 budget
    tol = 1e-6 if out["mode"] in ("fused32", "fusedbf16", "fusedbf16x3") else 1e-12
    if extra.get("layout") == "chain":
        assert out["layout"] == "chain"
    if extra.get("plan") == "auto":
        assert out["plan"]["source"] in ("cost model", "fallback")
    assert abs(out["norm2"] - 1) <= tol
    want = closed_form(name, prog)
    probs = np.abs(want) ** 2
    for q, p1 in out["p1"].items():
        q = int(q)
        exp1 = probs[((np.arange(probs.size) >> q) & 1) == 1].sum()
        assert abs(p1 - exp1) <= tol
    if out["measured"]:
        for shot in out["samples"]:
            assert all(bool((bv_mask(prog) >> q) & 1) == bool(b) for q, b in zip(out["measured"], shot))


@pytest.mark.gpu
@pytest.mark.parametrize("name,extra,check", [
    ("bv_n26", {"mode": "fusedbf16", "layout": "chain"}, "auto"),
    ("qft_n26", {"mode": "fusedbf16", "greedy": 7}, "auto"),
    ("qft_n26", {"mode": "fused32", "greedy": 5}, "fp64"),
    ("qf21_n15", {"nat": 8}, "inverse"),
    ("dnn_n16", {"mode": "fusedbf16", "greedy": 6}, "inverse"),
    ("qft_n26", {"mode": "fusedbf16x3", "layout": "chain"}, "auto"),
    ("cc_n18", {"mode": "fusedbf16x3", "greedy": 7}, "inverse"),
    ("qft_n26", {"mode": "fused64", "greedy": 6}, "inverse"),
])
def test_driver_accuracy_report(name, extra, check):
    """The driver's per-run accuracy report: against the FP64 per-gate path
    in a second state (the default for the split-precision modes), or by the
    inverse circuit back to |0...0>; the two-piece split paths stay within
    their 1e-6 infidelity budget, the three-piece split within 1e-10, the
    FP64 ones within 1e-12."""
    from hisvsim_amd import driver
    out = driver.run(_args(app=os.path.join(QDIR, name + ".qasm"), parts=True, verify=check, **extra))
    acc = out["accuracy"]
    assert acc is not None and acc["check"] == ("fp64" if check == "auto" else check)
    tol = 1e-6 if out["mode"] in ("fused32", "fusedbf16") else 1e-10 if out["mode"] == "fusedbf16x3" else 1e-12
    assert acc["infidelity"] <= tol
    if check == "inverse":
        assert acc["verified"] and acc["l2_err"] <= (1e-3 if tol == 1e-6 else 1e-5)
    else:
        assert acc["max_abs_err"] <= (1e-3 if tol == 1e-6 else 1e-6 if tol == 1e-10 else 1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("partfile", ["qft_n26_part_nat", "qft_n26_part_dfs", "qft_n26_part_smart"])
def test_qft26_shipped_partitions_on_gpu(partfile):
    """qft_n26 through each shipped partition (parts up to 26 qubits fall
    back to the per-gate path inside a part) equals the uniform state, the
    closed form of QFT|0>."""
    from hisvsim_amd import capi
    prog = prog_of("qft_n26")
    parts = qasm.parts_from_parts_file(prog, os.path.join(QDIR, partfile))
    with capi.State(26) as s:
        for p in parts:
            s.apply_part(p, 0)
        got = s.read()
    d, dn = phase_aligned_diff(got, closed_form("qft_n26", prog))
    assert d <= 1e-12 and dn <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["cc_n18", "dnn_n16", "qf21_n15", "sat_n11", "vqe_uccsd_n4", "vqe_uccsd_n8",
                                  "qpe_n9"])
def test_driver_circuits_match_oracle(orc, name):
    """Circuits without a closed form: GPU part path (natural-order parts,
    k = 10) against the CPU oracle's per-gate run, both lowering modes."""
    from hisvsim_amd import capi
    from hisvsim_amd.partition import nat_parts
    prog = prog_of(name)
    n = prog.num_qubits
    for mode in ("reference", "textbook"):
        gates = prog.gates(mode)
        want = np.zeros(2 ** n, complex)
        want[0] = 1
        orc.apply_gates(want, gates)
        with capi.State(n) as s:
            for p in nat_parts(gates, min(10, n)):
                s.apply_part(p, 0)
            got = s.read()
        assert np.abs(got - want).max() <= 1e-12 * max(1.0, np.abs(want).max())


def test_measured_qubits_and_deferral():
    from hisvsim_amd import driver
    p = qasm.parse("qreg q[3]; creg c[3]; h q[0]; measure q[0] -> c[0]; h q[1]; measure q[1] -> c[1];")
    assert driver.measured_qubits(p) == [0, 1]
    p = qasm.parse("qreg q[2]; creg c[2]; h q[0]; measure q[0] -> c[0]; x q[0];")
    with pytest.raises(qasm.QasmError):
        driver.measured_qubits(p)


def test_write_dot_files(tmp_path):
    """One DAG file per measurement-separated segment (qpe_n9 has six)."""
    prog = prog_of("qpe_n9")
    paths = qasm.write_dot_files(prog, str(tmp_path / "qpe_n9.qasm"))
    assert [os.path.basename(p) for p in paths] == [f"qpe_n9.qasm_PART_{i}.dot" for i in range(6)]
    for i, p in enumerate(paths):
        assert hashlib.sha256(open(p, "rb").read()).hexdigest() == FIX["dot"][f"qpe_n9.qasm_PART_{i}.dot"]["sha256"]
