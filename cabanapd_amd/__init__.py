"""MI355X-native peridynamics force/update engine (CabanaPD hot path).

The product is libpdk.so (C ABI, include/pdk.h) and the C++ drop-in headers
(include/CabanaPD/).  This package holds the engine sources (csrc/), the
ctypes binding used by the tests and the benchmark (engine.py) and the
workload generators (configs.py).
"""
import os
import subprocess

from . import _abi  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs=8):
    """Compile libpdk.so for gfx950 (always recompiles what changed)."""
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "csrc"), "-j%d" % jobs], check=True)


def library_path():
    return os.path.join(HERE, "libpdk.so")
