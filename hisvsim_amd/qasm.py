"""OpenQASM 2.0 front end: parser, gate lowering, the ordered-gate dump and
the gate-DAG (``.dot``) writer.

Counterpart of the reference driver qasm_assembler_standalone.py:
  * parsing and ordering -- ``QASM_parser`` (:64-409): qreg declaration order
    fixes qubit ids (:3569-3575, ``create_universal_qubit``); ``barrier`` is
    dropped, ``measure`` is kept in the ordered list (:209-281); a gate on a
    whole register is applied to each of its qubits (:283-297);
  * the ordered-gate dump ``<file>_exec`` (``dump_qasm``, :410-423);
  * gate lowering to simulator gates (``execute_qasm_parts``, :1803-2010);
  * the gate DAG per measurement-free segment, written as
    ``<file>_PART_<i>.dot`` (``produce_dot_file`` and ``add_input_node_pass``,
    :3619-3764), whose node names and order ids the parts files refer to.

Where the reference parser is line-based and quirky, this one is a statement
parser with textbook OpenQASM 2.0 semantics; the differences (documented in
the project's design notes) are: several statements per line are all read;
parameter expressions are evaluated exactly (``3*pi/4``, ``1.5e-3``,
functions); gate definitions with parameters are supported; gates before the
last gate definition are kept; ``measure q -> c`` on whole registers is
expanded per qubit; ``if`` and ``opaque`` are rejected with an error instead
of an abort. On the reference's own circuits the ordered list, the dump and
the DAG files are identical (the test suite compares them byte for byte).
"""
from __future__ import annotations

import ast
import math
import operator
import re
from collections import OrderedDict
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

from .circuit import Gate

# Gate names the reference accepts from a QASM file (qasm_assembler_standalone.py:19),
# plus the ones this front end adds (textbook qelib1 gates the reference lacks).
REFERENCE_GATES = ('u3', 'u2', 'u1', 'id', 'x', 'y', 'z', 'h', 's', 'sdg', 't', 'tdg', 'rx', 'ry', 'rz', 'cz',
                   'cx', 'cy', 'ch', 'ccx', 'crz', 'cu1', 'cu3', 'reset', 'swap', 'rxx', 'ryy', 'rzz', 'cswap')
EXTRA_GATES = ('u', 'p', 'cp', 'sx', 'sxdg', 'crx', 'cry', 'c3x', 'c4x', 'u0')
BUILTIN_GATES = set(REFERENCE_GATES) | set(EXTRA_GATES) | {'CX', 'U'}


class QasmError(ValueError):
    pass


@dataclass
class QasmOp:
    """One entry of the ordered gate list (the reference's ``TargetGate``,
    qasm_assembler_standalone.py:35-61): ``index`` is the 0-based source
    line of the statement (for a gate expanded from a ``gate`` definition,
    the line inside the definition), ``qubits`` are ``reg[i]`` strings (a
    measurement also lists its classical bit)."""
    index: int
    gate: str
    params: List[float]
    qubits: List[str]


@dataclass
class _GateDef:
    name: str
    params: List[str]
    args: List[str]
    body: List[Tuple[int, str, List[str], List[str]]]  # (line, name, param exprs, arg names)


@dataclass
class QasmProgram:
    qregs: "OrderedDict[str, int]" = field(default_factory=OrderedDict)
    cregs: "OrderedDict[str, int]" = field(default_factory=OrderedDict)
    ops: List[QasmOp] = field(default_factory=list)
    filename: str = ""

    @property
    def num_qubits(self) -> int:
        return sum(self.qregs.values())

    def qubit_ids(self) -> Dict[str, int]:
        """``reg[i]`` -> qubit id, registers in declaration order
        (qasm_assembler_standalone.py:3569-3575)."""
        ids: Dict[str, int] = {}
        for name, size in self.qregs.items():
            for i in range(size):
                ids[f"{name}[{i}]"] = len(ids)
        return ids

    def gates(self, mode: str = "reference") -> List[Gate]:
        """Simulator gates of the whole program (measurements skipped)."""
        ids = self.qubit_ids()
        return [g for g in (lower_op(op, ids, mode) for op in self.ops) if g is not None]


# ---- parameter expressions ------------------------------------------------------
_BINOPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.Div: operator.truediv,
           ast.Pow: operator.pow}
_FUNCS = {"sin": math.sin, "cos": math.cos, "tan": math.tan, "exp": math.exp, "ln": math.log, "sqrt": math.sqrt}


def eval_param(expr: str, env: Optional[Dict[str, float]] = None) -> float:
    """Evaluates an OpenQASM 2.0 parameter expression (numbers, ``pi``,
    + - * / ^, unary minus, sin cos tan exp ln sqrt, and the enclosing gate
    definition's parameters)."""
    env = env or {}
    try:
        tree = ast.parse(expr.strip().replace("^", "**"), mode="eval")
    except SyntaxError as e:
        raise QasmError(f"bad parameter expression {expr!r}") from e

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
            return float(n.value)
        if isinstance(n, ast.Name):
            if n.id == "pi":
                return math.pi
            if n.id in env:
                return env[n.id]
            raise QasmError(f"unknown name {n.id!r} in {expr!r}")
        if isinstance(n, ast.BinOp) and type(n.op) in _BINOPS:
            return _BINOPS[type(n.op)](ev(n.left), ev(n.right))
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd)):
            v = ev(n.operand)
            return -v if isinstance(n.op, ast.USub) else v
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in _FUNCS and len(n.args) == 1:
            return _FUNCS[n.func.id](ev(n.args[0]))
        raise QasmError(f"unsupported parameter expression {expr!r}")

    return float(ev(tree))


# ---- statement scanner --------------------------------------------------------------
def _statements(text: str):
    """Yields (line, statement text) with comments removed; a ``gate``
    definition is yielded whole, body included, as one statement."""
    lines = text.split("\n")
    clean = [ln.split("//", 1)[0] for ln in lines]
    buf, start, depth = [], None, 0
    for li, ln in enumerate(clean):
        for ch in ln:
            if start is None:
                if ch.isspace():
                    continue
                start = li
            buf.append(ch)
            if ch == "{":
                depth += 1
            elif ch == "}":
                depth -= 1
                if depth == 0:
                    yield start, "".join(buf).strip(), li
                    buf, start = [], None
            elif ch == ";" and depth == 0:
                yield start, "".join(buf)[:-1].strip(), li
                buf, start = [], None
        if start is not None:
            buf.append("\n")
    if start is not None and "".join(buf).strip():
        raise QasmError(f"unterminated statement at line {start + 1}")


_APPLY = re.compile(r"^([A-Za-z_][A-Za-z0-9_]*)\s*(?:\((.*)\))?\s*(.*)$", re.S)
_ARG = re.compile(r"^([A-Za-z_][A-Za-z0-9_]*)\s*(?:\[\s*(\d+)\s*\])?$")


def _split_params(s: str) -> List[str]:
    out, depth, cur = [], 0, []
    for ch in s:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            out.append("".join(cur))
            cur = []
        else:
            cur.append(ch)
    if "".join(cur).strip():
        out.append("".join(cur))
    return [p.strip() for p in out]


def _gate_body_lines(text: str, body_start_line: int):
    """Statements of a gate body with their source lines."""
    line = body_start_line
    cur, cur_line = [], None
    for ch in text:
        if ch == "\n":
            line += 1
            if cur:
                cur.append(" ")
            continue
        if cur_line is None:
            if ch.isspace():
                continue
            cur_line = line
        if ch == ";":
            yield cur_line, "".join(cur).strip()
            cur, cur_line = [], None
        else:
            cur.append(ch)
    if cur and "".join(cur).strip():
        yield cur_line, "".join(cur).strip()


def parse(text: str, filename: str = "") -> QasmProgram:
    prog = QasmProgram(filename=filename)
    gdefs: Dict[str, _GateDef] = {}

    def reg_args(argtext: str, line: int) -> List[Tuple[str, Optional[int]]]:
        args = []
        for a in [x.strip() for x in argtext.split(",") if x.strip()]:
            m = _ARG.match(a)
            if not m:
                raise QasmError(f"line {line + 1}: bad argument {a!r}")
            args.append((m.group(1), None if m.group(2) is None else int(m.group(2))))
        return args

    def emit(line: int, name: str, params: List[float], qubits: List[str]):
        if name in gdefs:
            d = gdefs[name]
            if len(params) != len(d.params) or len(qubits) != len(d.args):
                raise QasmError(f"line {line + 1}: gate {name} takes {len(d.params)} parameters and "
                                f"{len(d.args)} qubits")
            env = dict(zip(d.params, params))
            amap = dict(zip(d.args, qubits))
            for bl, bname, bparams, bargs in d.body:
                if bname == "barrier":
                    continue
                try:
                    emit(bl, bname, [eval_param(p, env) for p in bparams], [amap[a] for a in bargs])
                except KeyError as e:
                    raise QasmError(f"line {bl + 1}: {e.args[0]!r} is not an argument of gate {name}") from e
            return
        if name not in BUILTIN_GATES:
            raise QasmError(f"line {line + 1}: {name} not in function list")
        prog.ops.append(QasmOp(line, name, params, qubits))

    for line, st, end_line in _statements(text):
        head = st.split(None, 1)[0] if st else ""
        if not st or head in ("OPENQASM", "include"):
            continue
        if head in ("qreg", "creg"):
            m = re.match(r"^(qreg|creg)\s+([A-Za-z_][A-Za-z0-9_]*)\s*\[\s*(\d+)\s*\]$", st)
            if not m:
                raise QasmError(f"line {line + 1}: bad declaration {st!r}")
            (prog.qregs if m.group(1) == "qreg" else prog.cregs)[m.group(2)] = int(m.group(3))
            continue
        if head == "gate":
            m = re.match(r"^gate\s+([A-Za-z_][A-Za-z0-9_]*)\s*(?:\(([^)]*)\))?\s*([^{]*)\{(.*)\}$", st, re.S)
            if not m:
                raise QasmError(f"line {line + 1}: bad gate definition")
            name = m.group(1)
            params = [p.strip() for p in (m.group(2) or "").split(",") if p.strip()]
            args = [a.strip() for a in m.group(3).split(",") if a.strip()]
            # body lines: count the newlines before the opening brace
            pre = st[: st.index("{") + 1]
            body_line = line + pre.count("\n")
            body = []
            for bl, bst in _gate_body_lines(m.group(4), body_line):
                bm = _APPLY.match(bst)
                if not bm:
                    raise QasmError(f"line {bl + 1}: bad statement in gate {name}")
                body.append((bl, bm.group(1), _split_params(bm.group(2) or ""),
                             [a.strip() for a in bm.group(3).split(",") if a.strip()]))
            gdefs[name] = _GateDef(name, params, args, body)
            continue
        if head in ("if", "opaque"):
            raise QasmError(f"line {line + 1}: '{head}' is not supported")
        if head == "barrier":
            continue
        if head == "measure":
            m = re.match(r"^measure\s+(.+?)\s*->\s*(.+)$", st)
            if not m:
                raise QasmError(f"line {line + 1}: bad measure")
            (qr, qi), = reg_args(m.group(1), line)
            (cr, ci), = reg_args(m.group(2), line)
            if qi is None:
                for i in range(prog.qregs[qr]):
                    prog.ops.append(QasmOp(line, "measure", [], [f"{qr}[{i}]", f"{cr}[{i}]"]))
            else:
                prog.ops.append(QasmOp(line, "measure", [], [f"{qr}[{qi}]", f"{cr}[{ci}]"]))
            continue
        m = _APPLY.match(st)
        if not m:
            raise QasmError(f"line {line + 1}: cannot parse {st!r}")
        name = m.group(1)
        params = [eval_param(p) for p in _split_params(m.group(2) or "")]
        args = reg_args(m.group(3), line)
        for r, _ in args:
            if r not in prog.qregs:
                raise QasmError(f"line {line + 1}: unknown register {r!r}")
        sizes = {prog.qregs[r] for r, i in args if i is None}
        if len(sizes) > 1:
            raise QasmError(f"line {line + 1}: registers of different sizes")
        if sizes:  # broadcast over whole registers
            for j in range(sizes.pop()):
                emit(line, name, params, [f"{r}[{j if i is None else i}]" for r, i in args])
        else:
            emit(line, name, params, [f"{r}[{i}]" for r, i in args])
    return prog


def parse_file(path: str) -> QasmProgram:
    with open(path) as f:
        return parse(f.read(), filename=path)


# ---- the ordered-gate dump (qasm_assembler_standalone.py:410-423) -------------------
def _py_str(v: float) -> str:
    return str(v)


def exec_dump(prog: QasmProgram) -> str:
    out = ["QUBIT ALLOCATION"]
    for name, size in prog.qregs.items():
        out += [name, f"NUM: {size}"]
    for op in prog.ops:
        out += ["------", f"gate: {op.gate}", "parameters: ", " ".join(_py_str(p) for p in op.params),
                "on qubits", " ".join(op.qubits)]
    return "\n".join(out) + "\n"


# ---- lowering (qasm_assembler_standalone.py:1803-2010) ----------------------------
_ONE_Q = {"x": "X", "y": "Y", "z": "Z", "h": "H", "s": "S", "t": "T", "sdg": "SDG", "tdg": "TDG",
          "rx": "RX", "ry": "RY", "rz": "RZ", "u1": "U1", "u2": "U2", "u3": "U3", "id": "I", "u0": "I",
          "U": "U3", "u": "U3"}
_CONTROLLED = {"cx": "X", "CX": "X", "cy": "Y", "cz": "Z", "ch": "H", "crx": "RX", "cry": "RY", "crz": "RZ",
               "cu1": "U1", "cu3": "U3", "cp": "R1"}
_MULTI_X = {"ccx": 2, "c3x": 3, "c4x": 4}
_SX = (0.5 + 0.5j, 0.5 - 0.5j, 0.5 - 0.5j, 0.5 + 0.5j)


def _textbook_u3(theta: float, phi: float, lam: float) -> tuple:
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    e = lambda a: complex(math.cos(a), math.sin(a))  # noqa: E731
    return (c, -e(lam) * s, e(phi) * s, e(phi + lam) * c)


def lower_op(op: QasmOp, ids: Dict[str, int], mode: str = "reference") -> Optional[Gate]:
    """One ordered-list entry -> simulator gate (``None`` for a measurement).

    ``mode="reference"`` follows the reference lowering: ``u1``/``cu1`` are
    the reference's U1 (= RZ, basic_gates.hpp:571) and ``u2``/``u3``/``cu3``
    carry its phase convention. ``mode="textbook"`` uses the OpenQASM
    matrices: ``u1``/``cu1`` -> phase gate R1, ``u2``/``u3``/``cu3`` -> the
    textbook U3 matrix as a C1 gate. In both modes ``swap`` is a two-target
    SWAP, ``p``/``cp`` are (controlled) R1 and ``sx``/``sxdg`` are C1 gates,
    where the reference driver has no working lowering."""
    g = op.gate
    q = [ids[x] for x in op.qubits] if g != "measure" else []
    p = list(op.params)
    textbook = mode == "textbook"
    if mode not in ("reference", "textbook"):
        raise ValueError(f"unknown lowering mode {mode!r}")
    if g == "measure":
        return None
    if g == "reset":
        raise QasmError(f"line {op.index + 1}: reset is not supported by the state-vector path")

    def need(nq, np_):
        if len(q) != nq or len(p) != np_:
            raise QasmError(f"line {op.index + 1}: {g} takes {np_} parameters and {nq} qubits")

    if g in _ONE_Q:
        name = _ONE_Q[g]
        need(1, {"U1": 1, "U2": 2, "U3": 3, "RX": 1, "RY": 1, "RZ": 1}.get(name, 0))
        if textbook and name == "U1":
            return Gate("R1", (q[0],), (), (p[0],))
        if textbook and name in ("U2", "U3"):
            th, ph, la = (math.pi / 2, p[0], p[1]) if name == "U2" else tuple(p)
            return Gate("C1", (q[0],), (), _textbook_u3(th, ph, la))
        return Gate(name, (q[0],), (), tuple(p))
    if g == "p":
        need(1, 1)
        return Gate("R1", (q[0],), (), (p[0],))
    if g in ("sx", "sxdg"):
        need(1, 0)
        m = _SX if g == "sx" else tuple(v.conjugate() for v in _SX)
        return Gate("C1", (q[0],), (), m)
    if g in _CONTROLLED:
        name = _CONTROLLED[g]
        need(2, {"RX": 1, "RY": 1, "RZ": 1, "U1": 1, "U3": 3, "R1": 1}.get(name, 0))
        if textbook and name == "U1":
            return Gate("R1", (q[1],), (q[0],), (p[0],))
        if textbook and name == "U3":
            return Gate("C1", (q[1],), (q[0],), _textbook_u3(*p))
        return Gate(name, (q[1],), (q[0],), tuple(p))
    if g in ("rxx", "ryy", "rzz"):
        need(2, 1)
        return Gate(g.upper(), (q[0], q[1]), (), (p[0],))
    if g == "swap":
        need(2, 0)
        return Gate("SWAP", (q[0], q[1]))
    if g == "cswap":
        need(3, 0)
        return Gate("SWAP", (q[1], q[2]), (q[0],))
    if g in _MULTI_X:
        c = _MULTI_X[g]
        need(c + 1, 0)
        return Gate("X", (q[c],), tuple(q[:c]))
    raise QasmError(f"line {op.index + 1}: no lowering for gate {g!r}")


# ---- gate DAG per segment (qasm_assembler_standalone.py:3619-3764) -----------------
class GateDag:
    """A multigraph with ordered nodes and edges, enough to reproduce the
    reference's DAG files: nodes keep insertion order, each node's out-edges
    are grouped by successor in first-insertion order, and parallel edges
    are numbered 0, 1, ... (the edge ``key``)."""

    def __init__(self):
        self.nodes: "OrderedDict[str, list]" = OrderedDict()
        self.succ: Dict[str, "OrderedDict[str, list]"] = {}
        self.pred: Dict[str, "OrderedDict[str, list]"] = {}

    def add_node(self, name: str, **attrs):
        if name in self.nodes:
            self.nodes[name] = list(dict(self.nodes[name], **attrs).items())
            return
        self.nodes[name] = list(attrs.items())
        self.succ[name] = OrderedDict()
        self.pred[name] = OrderedDict()

    def add_edge(self, u: str, v: str, label: str):
        self.succ[u].setdefault(v, []).append(label)
        self.pred[v].setdefault(u, []).append(label)

    def remove_node(self, x: str):
        for u in self.pred[x]:
            del self.succ[u][x]
        for v in self.succ[x]:
            del self.pred[v][x]
        del self.succ[x], self.pred[x], self.nodes[x]

    def in_degree(self, x):
        return sum(len(v) for v in self.pred[x].values())

    def out_degree(self, x):
        return sum(len(v) for v in self.succ[x].values())

    def edges(self):
        for u in self.nodes:
            for v, labels in self.succ[u].items():
                for key, lab in enumerate(labels):
                    yield u, v, key, lab

    def attr(self, name: str, key: str):
        return dict(self.nodes[name]).get(key)

    def to_dot(self) -> str:
        out = ["digraph  {"]
        for name, attrs in self.nodes.items():
            out.append(f"{_dot_id(name)} [{', '.join(f'{k}={_dot_id(str(v))}' for k, v in attrs)}];")
        for u, v, key, lab in self.edges():
            out.append(f"{_dot_id(u)} -> {_dot_id(v)}  [key={key}, label={_dot_id(lab)}];")
        out.append("}")
        return "\n".join(out) + "\n"


_DOT_PLAIN = re.compile(r"^([A-Za-z_][A-Za-z0-9_]*|-?(\.[0-9]+|[0-9]+(\.[0-9]*)?))$")
_DOT_KEYWORDS = {"node", "edge", "graph", "digraph", "subgraph", "strict"}


def _dot_id(s: str) -> str:
    if _DOT_PLAIN.match(s) and s.lower() not in _DOT_KEYWORDS:
        return s
    return '"' + s.replace('"', '\\"') + '"'


def _add_input_nodes(G: GateDag) -> None:
    """Adds a source node per qubit in front of the qubit's first gate
    (``add_input_node_pass``, qasm_assembler_standalone.py:3625-3670)."""
    need: "OrderedDict[str, list]" = OrderedDict()
    for node in list(G.nodes):
        if G.in_degree(node) == G.out_degree(node):
            continue
        in_labels = [lab for labs in G.pred[node].values() for lab in labs]
        for v, labs in G.succ[node].items():
            for lab in labs:
                if lab != "exit" and lab not in in_labels:
                    need.setdefault(node, []).append(lab)
    for node, qubits in need.items():
        for qb in qubits:
            G.add_node(qb, label=qb, order=0)
            G.add_edge(qb, node, qb)
    for node in G.nodes:
        if G.in_degree(node) == 0 or G.out_degree(node) == 0:
            continue
        ins = sorted(lab for labs in G.pred[node].values() for lab in labs)
        outs = sorted(lab for labs in G.succ[node].values() for lab in labs)
        if ins != outs:
            raise QasmError(f"DAG node {node} has in-edges {ins} but out-edges {outs}")


@dataclass
class DagSegment:
    dag: GateDag
    measures: List[QasmOp]   # measurements met while the segment was empty


def build_dags(ops: Sequence[QasmOp]) -> List[DagSegment]:
    """The gate DAG of each measurement-free segment of the ordered list.

    Node names are ``<gate>_<c>`` and ``<qubit>_exit_<c>`` with one counter
    ``c`` over the whole list that advances once per gate and once per exit
    node; a gate node carries its source line and ``order`` = position in
    the list + 1, exit nodes ``order`` = len(list) + 8192, input nodes (one
    per qubit, named after it) ``order`` = 0. A measurement closes the
    current segment; one met while the segment is empty is attached to it."""
    segments: List[DagSegment] = []
    G, live, pending = GateDag(), [], []   # live: exit nodes in creation order
    counter = 0
    exit_order = len(ops) + 8192
    for order, op in enumerate(ops):
        if op.gate == "measure":
            if not G.nodes:
                pending.append(op)
                continue
            _add_input_nodes(G)
            segments.append(DagSegment(G, pending))
            G, live, pending = GateDag(), [], []
            continue
        name = f"{op.gate}_{counter}"
        counter += 1
        G.add_node(name, label=op.gate, line=op.index, order=order + 1)
        labels = [qb.replace("[", "").replace("]", "") for qb in op.qubits]
        fresh = []
        for lab in labels:
            void = f"{lab}_exit_{counter}"
            counter += 1
            G.add_node(void, label="exit", order=exit_order)
            fresh.append(void)
            G.add_edge(name, void, lab)
        remaining = len(labels)
        dead = []
        for node in live:
            for u, labs in list(G.pred[node].items()):
                for lab in labs:
                    if lab in labels:
                        G.add_edge(u, name, lab)
                        dead.append(node)
                        remaining -= 1
            if remaining == 0:
                break
        for node in dead:
            G.remove_node(node)
            live.remove(node)
        live.extend(fresh)
    if G.nodes:
        _add_input_nodes(G)
        segments.append(DagSegment(G, pending))
    elif pending:
        segments.append(DagSegment(G, pending))
    return segments


def write_dot_files(prog: QasmProgram, prefix: Optional[str] = None) -> List[str]:
    """Writes ``<prefix>_PART_<i>.dot`` for each segment with gates; returns
    the paths."""
    prefix = prefix if prefix is not None else prog.filename
    paths = []
    for i, seg in enumerate(s for s in build_dags(prog.ops) if s.dag.nodes):
        path = f"{prefix}_PART_{i}.dot"
        with open(path, "w") as f:
            f.write(seg.dag.to_dot())
        paths.append(path)
    return paths


# ---- parts files (qasm_assembler_standalone.py:3549-3567) --------------------------
def parts_file_text(seg: DagSegment, part_of_order: Dict[int, int]) -> str:
    """A parts file for one DAG segment: ``"<order> <node> <part>"`` per
    node in DAG order. ``part_of_order`` gives each gate's part by its order
    id; an input node takes the part of the gate after it and an exit node
    the part of the gate before it (the convention of the reference's
    shipped parts files)."""
    G = seg.dag
    gate_nodes = part_of_order_names(G, part_of_order)
    out = []
    for name in G.nodes:
        order = G.attr(name, "order")
        if name in gate_nodes:
            part = part_of_order[order]
        elif order == 0:
            part = part_of_order[G.attr(next(iter(G.succ[name])), "order")]
        else:
            part = part_of_order[G.attr(next(iter(G.pred[name])), "order")]
        out.append(f"{order} {name} {part}")
    return "\n".join(out) + "\n"


def part_of_order_names(G: GateDag, part_of_order: Dict[int, int]) -> set:
    return {n for n in G.nodes if G.attr(n, "line") is not None and G.attr(n, "order") in part_of_order}


def gate_order_ids(prog: QasmProgram) -> List[int]:
    """Order id (position in the ordered list + 1) of every gate, skipping
    measurements: the ids parts files use for gates."""
    return [i + 1 for i, op in enumerate(prog.ops) if op.gate != "measure"]


def parts_from_parts_file(prog: QasmProgram, path: str, mode: str = "reference") -> List[List[Gate]]:
    """Simulator gates per part, in ascending part id and file order within
    a part; input nodes (order 0), exit nodes (order past the list) and
    measurements are skipped (qasm_assembler_standalone.py:1783-1798)."""
    from .partition import read_parts_file
    ids = prog.qubit_ids()
    out = []
    for _, orders in read_parts_file(path).items():
        part = []
        for o in orders:
            if o == 0 or o > len(prog.ops):
                continue
            g = lower_op(prog.ops[o - 1], ids, mode)
            if g is not None:
                part.append(g)
        if part:
            out.append(part)
    return out
