#!/usr/bin/env python3
"""Collects the reference's QASM fixtures into tests/golden/qasm/.

Copies, from QASMBench/cluster/<circuit>/ of the reference tree:
  * the ``.qasm`` circuit of every directory that holds a DAG file
    (``<circuit>.qasm_PART_<i>.dot``), an ordered-gate dump
    (``<circuit>.qasm_exec``) or a parts file (``<circuit>_part_<method>``);
  * the parts files themselves (inputs of the partitioned runs);
and records the SHA-256 and size of every ``.dot`` and ``_exec`` file in
``qasm_fixtures.json``, so the tests can compare this package's output with
the reference's byte for byte without carrying the larger files.

    python tests/golden/make_qasm_fixtures.py [REFERENCE_ROOT]
"""
import glob
import hashlib
import json
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "qasm")


def main() -> int:
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "QASMBench", "cluster")
    os.makedirs(OUT, exist_ok=True)
    table = {"dot": {}, "exec": {}, "parts": []}
    circuits = set()
    for path in sorted(glob.glob(os.path.join(src, "*", "*"))):
        base = os.path.basename(path)
        d = os.path.basename(os.path.dirname(path))
        if ".qasm_PART_" in base and base.endswith(".dot"):
            kind = "dot"
        elif base.endswith(".qasm_exec"):
            kind = "exec"
        elif "_part_" in base and not base.endswith(".txt"):
            kind = "parts"
        else:
            continue
        circuit = base.split(".qasm")[0] if kind != "parts" else base.rsplit("_part_", 1)[0]
        qasm = os.path.join(os.path.dirname(path), circuit + ".qasm")
        if not os.path.exists(qasm):
            continue
        circuits.add((d, circuit))
        if kind == "parts":
            shutil.copyfile(path, os.path.join(OUT, base))
            table["parts"].append(base)
        else:
            data = open(path, "rb").read()
            table[kind][base] = {"sha256": hashlib.sha256(data).hexdigest(), "bytes": len(data)}
    for d, circuit in sorted(circuits):
        shutil.copyfile(os.path.join(src, d, circuit + ".qasm"), os.path.join(OUT, circuit + ".qasm"))
    table["circuits"] = sorted(c for _, c in circuits)
    with open(os.path.join(OUT, "qasm_fixtures.json"), "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print(f"{len(circuits)} circuits, {len(table['dot'])} dot, {len(table['exec'])} exec, "
          f"{len(table['parts'])} parts files")
    return 0


if __name__ == "__main__":
    sys.exit(main())
