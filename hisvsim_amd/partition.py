"""Acyclic partitioners and the parts-file format.

The reference takes its parts from an external acyclic DAG partitioner (dagP,
not vendored) and reads them from a parts file with one line per DAG node,
``"<order> <node> <part>"`` (qasm_assembler_standalone.py:3549-3567); parts run
in ascending part id, gates within a part in file order, order 0 (the input
node) and the exit nodes are skipped (:1783-1791).

Partitioners here (all return a list of parts, each a list of gates, in
execution order; every part touches at most k qubits and the sequence is a
valid topological order of the gate DAG):

* ``nat_parts`` -- natural-order greedy: start a new part when the next gate
  would push the part past k qubits. Reproduces the shipped ``*_part_nat``.
* ``greedy_parts`` -- DAG-aware greedy: fills each part by scanning the
  remaining gates in order, taking every gate whose qubits fit and whose
  predecessors are all placed; a skipped gate blocks its qubits for the rest
  of the scan, which keeps the quotient DAG acyclic.
* ``layered_parts`` -- for circuits made of single-qubit layers and two-qubit
  matchings (the random layered benchmark): groups each matching's pairs so a
  part holds whole pairs plus the single-qubit gates that follow them.
* ``dfs_parts`` -- natural-order parts over a depth-first topological order
  of the gate DAG (the method of the shipped ``*_part_dfs`` files, with this
  package's own DFS).
* ``merge_parts`` -- refinement: joins neighbouring parts whose qubits fit.
"""
from __future__ import annotations

import heapq
from collections import OrderedDict
from typing import Dict, List, Sequence, Tuple

from .circuit import Gate


def gate_qubits(g: Gate) -> set:
    return set(g.controls) | set(g.targets)


def tile_qubits(g: Gate) -> set:
    """The qubits a gate needs in a BF16-split tile: none for an identity
    (hsv_apply_fused_chain leaves I gates out of a part's tile, so a qubit
    that only an I touches must not count towards the tile or be offered as
    a next low slot)."""
    return set() if g.name.upper() == "I" else gate_qubits(g)


def nat_parts(gates: Sequence[Gate], k: int, always: Sequence[int] = ()) -> List[List[Gate]]:
    """``always`` as in ``greedy_parts``: qubits counted in every part."""
    base = set(always)
    if len(base) > k:
        raise ValueError("more always-included qubits than the part width")
    parts: List[List[Gate]] = []
    cur: List[Gate] = []
    qs: set = set(base)
    for g in gates:
        gq = gate_qubits(g)
        if len(base | gq) > k:
            raise ValueError(f"gate {g} does not fit a part of {k} qubits that already holds {sorted(base)}")
        if cur and len(qs | gq) > k:
            parts.append(cur)
            cur, qs = [], set(base)
        cur.append(g)
        qs |= gq
    if cur:
        parts.append(cur)
    return parts


def nat_part_ids(gates: Sequence[Gate], k: int) -> List[int]:
    """Part id of every gate under ``nat_parts`` (ids start at 0)."""
    ids, pid, qs = [], 0, set()
    for g in gates:
        gq = gate_qubits(g)
        if ids and len(qs | gq) > k:
            pid += 1
            qs = set()
        qs |= gq
        ids.append(pid)
    return ids


def greedy_parts(gates: Sequence[Gate], k: int, always: Sequence[int] = ()) -> List[List[Gate]]:
    """``always``: qubits every part includes anyway (the device adds the
    lowest qubits to each part for contiguous HBM runs); they count towards k
    from the start, so gates on them join parts for free."""
    remaining = list(range(len(gates)))
    parts: List[List[Gate]] = []
    base_q = set(always)
    if len(base_q) > k:
        raise ValueError("more always-included qubits than the part width")
    while remaining:
        part_q: set = set(base_q)
        blocked: set = set()
        taken: List[int] = []
        keep: List[int] = []
        for i in remaining:
            gq = gate_qubits(gates[i])
            if gq & blocked:
                blocked |= gq
                keep.append(i)
            elif len(part_q | gq) <= k:
                part_q |= gq
                taken.append(i)
            else:
                blocked |= gq
                keep.append(i)
        if not taken:
            raise ValueError(f"gate {gates[remaining[0]]} does not fit a part of {k} qubits "
                             f"that already holds {sorted(base_q)}")
        parts.append([gates[i] for i in taken])
        remaining = keep
    return parts


def layered_parts(gates: Sequence[Gate], n: int, k: int) -> List[List[Gate]]:
    """Parts for layer-structured circuits: each layer is single-qubit gates on
    some qubits followed by two-qubit gates on disjoint pairs.

    The circuit is cut into segments at the start of every two-qubit block;
    segment L holds the pairs of block L and the single-qubit gates after them
    (up to the next block). Pairs are grouped k//2 at a time (an idle qubit
    fills a slot if k is odd or a group is short), and each group's part holds
    its pairs plus the following single-qubit gates on its qubits. Gates
    before the first block are merged into the first segment's parts.
    Falls back to ``greedy_parts`` if the circuit does not have this shape.
    """
    # split into alternating runs of single-qubit and two-qubit gates
    runs: List[tuple] = []
    for g in gates:
        two = len(gate_qubits(g)) == 2
        if runs and runs[-1][0] == two:
            runs[-1][1].append(g)
        else:
            runs.append((two, [g]))
    if any(len(gate_qubits(g)) > 2 for g in gates):
        return greedy_parts(gates, k)
    for two, rg in runs:
        if two:
            seen: set = set()
            for g in rg:
                if gate_qubits(g) & seen:
                    return greedy_parts(gates, k)
                seen |= gate_qubits(g)
    pre: List[Gate] = []
    segments: List[tuple] = []  # (pairs gates, following single-qubit gates)
    i = 0
    if runs and not runs[0][0]:
        pre = runs[0][1]
        i = 1
    while i < len(runs):
        pairs = runs[i][1]
        singles = runs[i + 1][1] if i + 1 < len(runs) else []
        segments.append((pairs, singles))
        i += 2
    if not segments:
        return greedy_parts(gates, k)
    parts: List[List[Gate]] = []
    for si, (pairs, singles) in enumerate(segments):
        paired = set()
        for g in pairs:
            paired |= gate_qubits(g)
        idle = [q for q in range(n) if q not in paired]
        per = max(1, k // 2)
        groups: List[tuple] = []
        for j in range(0, len(pairs), per):
            groups.append((pairs[j:j + per], set().union(*[gate_qubits(g) for g in pairs[j:j + per]])))
        # idle qubits join groups with room, else form their own
        for q in idle:
            for gi, (pg, qs) in enumerate(groups):
                if len(qs) < k:
                    qs.add(q)
                    break
            else:
                groups.append(([], {q}))
        by_qubit: Dict[int, List[Gate]] = {}
        for g in singles:
            by_qubit.setdefault(g.targets[0], []).append(g)
        pre_by_q: Dict[int, List[Gate]] = {}
        if si == 0:
            for g in pre:
                pre_by_q.setdefault(g.targets[0], []).append(g)
        for pg, qs in groups:
            part: List[Gate] = []
            for q in sorted(qs):
                part += pre_by_q.get(q, [])
            part += pg
            for q in sorted(qs):
                part += by_qubit.get(q, [])
            if part:
                parts.append(part)
    return parts


def gate_dag(gates: Sequence[Gate]) -> List[List[int]]:
    """Successor lists of the gate DAG: an edge from each gate to the next
    gate on each of its qubits (the edges of the reference's DAG files,
    qasm_assembler_standalone.py:3688-3764, without input/exit nodes)."""
    succ: List[List[int]] = [[] for _ in gates]
    last: Dict[int, int] = {}
    for i, g in enumerate(gates):
        for q in sorted(gate_qubits(g)):
            if q in last and i not in succ[last[q]]:
                succ[last[q]].append(i)
            last[q] = i
    return succ


def dfs_order(gates: Sequence[Gate]) -> List[int]:
    """A topological order of the gate DAG that follows one dependency chain
    as deep as it can before backtracking (depth-first): among the ready
    gates, take a successor of the gate just placed when one is ready, else
    the earliest ready gate. The reference's shipped ``*_part_dfs`` files
    come from a DFS ordering of the same DAG by an external tool; this is
    this package's own."""
    succ = gate_dag(gates)
    indeg = [0] * len(gates)
    for s in succ:
        for j in s:
            indeg[j] += 1
    ready = sorted(i for i in range(len(gates)) if indeg[i] == 0)
    stack: List[int] = []
    order: List[int] = []
    heapq.heapify(ready)
    placed = [False] * len(gates)
    while len(order) < len(gates):
        nxt = None
        while stack and nxt is None:
            cand = [j for j in succ[stack[-1]] if indeg[j] == 0 and not placed[j]]
            if cand:
                nxt = min(cand)
            else:
                stack.pop()
        if nxt is None:
            while placed[ready[0]]:
                heapq.heappop(ready)
            nxt = heapq.heappop(ready)
        placed[nxt] = True
        order.append(nxt)
        stack.append(nxt)
        for j in succ[nxt]:
            indeg[j] -= 1
            if indeg[j] == 0:
                heapq.heappush(ready, j)
    return order


def dfs_parts(gates: Sequence[Gate], k: int, always: Sequence[int] = ()) -> List[List[Gate]]:
    """Natural-order parts over the depth-first topological order."""
    return nat_parts([gates[i] for i in dfs_order(gates)], k, always)


def _absorb(qsets: Sequence[frozenset], remaining: Sequence[int], Q: set) -> List[int]:
    """The blocking scan of ``greedy_parts`` for a fixed qubit set Q: the
    remaining gates, in order, that act inside Q and follow no skipped gate
    on a shared qubit."""
    blocked: set = set()
    taken: List[int] = []
    for i in remaining:
        gq = qsets[i]
        if gq & blocked or not gq <= Q:
            blocked |= gq
            if Q <= blocked:
                break
            continue
        taken.append(i)
    return taken


def closure_parts(gates: Sequence[Gate], k: int, always: Sequence[int] = (),
                  candidates: int = 24) -> List[List[Gate]]:
    """Qubit-set-growing clusterer. Each part starts from the ``always``
    qubits and grows its qubit set Q step by step: a step looks at the
    qubits that the first ``candidates`` remaining gates not yet inside Q
    (and not behind a skipped gate) would add, and adds the set with the most
    absorbed gates per added qubit, a gate being absorbed when the blocking
    scan of ``greedy_parts`` for Q takes it. The part is that scan's result
    for the final Q, so the sequence of parts is acyclic for the same reason
    as ``greedy_parts``. On layered circuits this packs markedly more gates
    per part than the scan-order methods (an acyclic clusterer in the role
    of the reference's external dagP partitioner, whose method is not
    shipped)."""
    qsets = [frozenset(gate_qubits(g)) for g in gates]
    base = set(always)
    if len(base) > k:
        raise ValueError("more always-included qubits than the part width")
    remaining = list(range(len(gates)))
    parts: List[List[Gate]] = []
    while remaining:
        taken = _absorb(qsets, remaining, _grow(qsets, remaining, base, k, candidates))
        if not taken:
            raise ValueError(f"gate {gates[remaining[0]]} does not fit a part of {k} qubits "
                             f"that already holds {sorted(base)}")
        done = set(taken)
        parts.append([gates[i] for i in taken])
        remaining = [i for i in remaining if i not in done]
    return parts


def _grow(qsets: Sequence[frozenset], remaining: Sequence[int], base: set, k: int,
          candidates: int = 24) -> set:
    """The qubit-set growth of ``closure_parts`` for one part, from ``base``."""
    Q = set(base)
    while len(Q) < k:
        blocked: set = set()
        cands: List[frozenset] = []
        for i in remaining:
            gq = qsets[i]
            if gq <= Q:
                continue
            add = gq - Q
            if not (gq & blocked) and len(Q) + len(add) <= k and add not in cands:
                cands.append(add)
            blocked |= gq
            if len(cands) >= candidates:
                break
        if not cands:
            break
        best, best_score = None, -1.0
        for add in cands:
            score = len(_absorb(qsets, remaining, Q | add)) / len(add)
            if score > best_score:
                best, best_score = add, score
        Q |= best
    return Q


def chain_parts(gates: Sequence[Gate], k: int, low: int = 5, lookahead: int = 13,
                candidates: int = 24, native: bool = True,
                beam: Tuple[int, int] = (1, 1)) -> Tuple[List[List[Gate]], List[List[int]]]:
    """Parts for a moving qubit layout (hsv_apply_fused_chain): part p is
    grown by the ``closure_parts`` method from the ``low`` qubits that sit in
    the lowest slots while it runs (lows[p]; lows[0] = qubits 0..low-1), so
    its tile - its qubits and lows[p] - has at most k qubits. lows[p + 1] is
    chosen among part p's tile: of the ``lookahead`` qubits whose next gates
    come first, the ``low`` that let the next part absorb the most gates
    (lookahead 13 tries every 5-subset of a 13-qubit tile). Unlike fixed low
    qubits, the low slots then hold qubits the next part uses. Returns
    (parts, lows); the parts are acyclic as for closure_parts. With
    ``native`` the same method runs in libhsv (hsv_chain_parts, ~80x faster,
    identical result) when the library loads. ``beam`` = (width, branch)
    widens the choice of lows into a beam search (native only: each chain in
    flight branches into its ``branch`` best low sets, the ``width`` with the
    fewest gates left go on; the shorter of that and the greedy chain is
    returned); (1, 1) is the greedy chain above."""
    if native:
        try:
            from . import capi
            capi.lib()
        except Exception:   # no library: the Python method below
            native = False
        else:
            if low > k:
                raise ValueError("more low qubits than the part width")
            return capi.chain_parts(gates, k, low, max(lookahead, low), candidates, beam)
    if tuple(beam) != (1, 1):
        raise ValueError("the beam search over low sets runs in libhsv only")
    from itertools import combinations
    if low > k:
        raise ValueError("more low qubits than the part width")
    qsets = [frozenset(tile_qubits(g)) for g in gates]
    nq = max((max(q) for q in qsets if q), default=-1) + 1
    if nq < low:
        raise ValueError(f"the circuit has {nq} qubits: fewer than the {low} low slots")
    remaining = list(range(len(gates)))
    lowset = set(range(low))
    parts: List[List[Gate]] = []
    lows: List[List[int]] = []
    while remaining:
        taken = _absorb(qsets, remaining, _grow(qsets, remaining, lowset, k, candidates))
        if not taken:
            raise ValueError(f"gate {gates[remaining[0]]} does not fit a part of {k} qubits "
                             f"that already holds {sorted(lowset)}")
        done = set(taken)
        parts.append([gates[i] for i in taken])
        lows.append(sorted(lowset))
        tile = set(lowset).union(*[qsets[i] for i in taken])
        remaining = [i for i in remaining if i not in done]
        if not remaining:
            break
        first: dict = {}
        for pos, i in enumerate(remaining):
            for q in qsets[i]:
                if q in tile and q not in first:
                    first[q] = pos
        order = sorted(tile, key=lambda q: (first.get(q, len(remaining)), q))
        best = None
        for cand in combinations(order[:max(lookahead, low)], low):
            got = len(_absorb(qsets, remaining, _grow(qsets, remaining, set(cand), k, candidates)))
            if best is None or got > best[0]:
                best = (got, set(cand))
        lowset = best[1]
    return parts, lows


def merge_parts(parts: Sequence[Sequence[Gate]], k: int) -> List[List[Gate]]:
    """Refinement: joins each part with the next while their qubit union
    stays within k. Joining neighbours in the sequence keeps every qubit's
    gate order, so any valid sequence stays valid."""
    out: List[List[Gate]] = []
    qs: set = set()
    for p in parts:
        pq = set().union(*[gate_qubits(g) for g in p]) if p else set()
        if out and len(qs | pq) <= k:
            out[-1] = out[-1] + list(p)
            qs |= pq
        else:
            out.append(list(p))
            qs = set(pq)
    return out


def check_parts(gates: Sequence[Gate], parts: Sequence[Sequence[Gate]], k: int | None = None,
                identities_free: bool = False) -> None:
    """Raises ValueError unless ``parts`` is a k-bounded acyclic partition of
    ``gates`` whose sequence respects every qubit's gate order. With
    ``identities_free`` (the chain planner's view) I gates occupy no qubit:
    they neither count towards k nor order anything."""
    qubits_of = tile_qubits if identities_free else gate_qubits
    pos = {}
    for i, g in enumerate(gates):
        pos.setdefault(g, []).append(i)
    flat = [g for p in parts for g in p]
    if len(flat) != len(gates):
        raise ValueError("parts do not cover the gates exactly once")
    # each qubit's gate sequence must be unchanged
    def per_qubit(seq):
        d: Dict[int, list] = {}
        for g in seq:
            for q in qubits_of(g):
                d.setdefault(q, []).append(g)
        return d
    if per_qubit(gates) != per_qubit(flat):
        raise ValueError("parts reorder gates on some qubit")
    if k is not None:
        for p in parts:
            if len(set().union(*[qubits_of(g) for g in p])) > k:
                raise ValueError("a part exceeds k qubits")


# ---- parts-file format (qasm_assembler_standalone.py:3549-3567) -------------
def read_parts_file(path: str) -> "OrderedDict[int, List[int]]":
    """part id -> [order ids] in file order, parts sorted by id."""
    parts: Dict[int, List[int]] = {}
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if not line:
                continue
            items = line.split(" ")
            if len(items) != 3:
                raise ValueError(f"bad parts-file line {line!r}")
            order_id, part_id = int(items[0]), int(items[2])
            parts.setdefault(part_id, []).append(order_id)
    return OrderedDict(sorted(parts.items()))


def parts_from_file(path: str, ordered_gates: Sequence[Gate]) -> List[List[Gate]]:
    """Gate lists per part: order 0 (input node) and order ids past the gate
    list (exit nodes) are skipped; order o maps to ordered_gates[o - 1]."""
    out = []
    for _, orders in read_parts_file(path).items():
        part = [ordered_gates[o - 1] for o in orders if 0 < o <= len(ordered_gates)]
        if part:
            out.append(part)
    return out


def write_parts_file(path: str, node_names: Sequence[str], part_ids: Sequence[int]) -> None:
    """One ``"<order> <node> <part>"`` line per gate (orders from 1)."""
    with open(path, "w") as f:
        for i, (name, pid) in enumerate(zip(node_names, part_ids)):
            f.write(f"{i + 1} {name} {pid}\n")
