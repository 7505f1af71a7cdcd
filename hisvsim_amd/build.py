"""Builds libhsv.so (HIP, gfx950) and the svsimulator_py_wrapper pybind module in-tree.

    python -m hisvsim_amd.build          # or: python hisvsim_amd/build.py

Always recompiles. Objects go to hisvsim_amd/_build/ (git-ignored); the two
shared libraries land next to this file so they travel with the source tree.
"""
from __future__ import annotations

import os
import shlex
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HSV_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops keeps the compiler from emitting packed FP32 VALU ops.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function", "-Wno-unused-value",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", os.path.join(ROOT, "include")]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-I", os.path.join(ROOT, "include")]

HIP_SOURCES = ["hsv.hip", "fused.hip", "qaoa.hip", "mlp.hip", "batch.hip"]
CXX_SOURCES = ["gates.cpp", "partition.cpp"]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> dict:
    os.makedirs(OUT, exist_ok=True)
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OUT, src + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(OUT, src + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    lib = os.path.join(HERE, "libhsv.so")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-pthread", "-o", lib, *objs])

    import pybind11  # build-time only
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    mod = os.path.join(HERE, "svsimulator_py_wrapper" + ext)
    py_inc = sysconfig.get_paths()["include"]
    _run(["g++", *CXX_FLAGS, "-shared", "-I", pybind11.get_include(), "-I", py_inc,
          os.path.join(CSRC, "py_wrapper.cpp"), "-o", mod, "-L", HERE, "-lhsv",
          "-Wl,-rpath,$ORIGIN", "-Wl,--no-as-needed"])
    return {"libhsv": lib, "module": mod}


if __name__ == "__main__":
    out = build()
    print(out)
    sys.exit(0)
