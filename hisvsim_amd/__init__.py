"""hisvsim_amd -- MI355X-native hierarchical state-vector simulation.

The compute path is libhsv.so (HIP kernels for gfx950 behind the C ABI in
include/hsv.h). Two Python front ends sit on it:

* ``hisvsim_amd.svsimulator_py_wrapper`` -- the reference's pybind module API
  (StateVector, AggregateGate_*, BoundGate, gather_and_execute_on, ...);
* ``hisvsim_amd.capi`` -- a thin ctypes binding of the C ABI.

Build both in place with ``python -m hisvsim_amd.build``.
"""
import importlib
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def load_module():
    """Imports the in-tree pybind module (raises ImportError if not built)."""
    if HERE not in sys.path:
        sys.path.insert(0, HERE)
    return importlib.import_module("svsimulator_py_wrapper")


__all__ = ["load_module"]
