"""graphine_amd: GRE's scatter -> combine -> apply superstep on MI355X (gfx950).

The product is the HIP library ``libgre_dev.so`` (C ABI in ``include/gre_dev.h``)
and the GRE-compatible C++ host headers in ``graphine_amd/include``.  This
Python package only locates and builds them; ``graphine_amd.native`` is a ctypes
binding used by tests and the benchmark.
"""
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
INCLUDE_DIR = os.path.join(PKG_DIR, "include")
LIB_PATH = os.path.join(PKG_DIR, "libgre_dev.so")


def build(jobs=8):
    """Compile libgre_dev.so for gfx950 in-tree (always recompiles changed sources)."""
    subprocess.run(["make", "-s", "-C", PKG_DIR, f"-j{jobs}"], check=True)
    return LIB_PATH
