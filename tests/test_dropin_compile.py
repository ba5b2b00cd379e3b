"""Drop-in check (CPU): GRE's own drivers (usr/test_*.cpp) and vertex programs
(include/Kernels/*.hpp) compile and link UNCHANGED against this build's C++
headers (graphine_amd/include) and libgre_dev.so.

Runs only where the GRE source tree is present (/root/reference); nothing is
copied: the compiler reads the files where they lie, and only the Kernels
directory is put on the include path, so every other GRE header the drivers
name resolves to this build's headers.  The binaries stay in pytest's tmp dir.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GRE = "/root/reference"
DRIVERS = ["pagerank", "bfs", "cc", "sssp", "bc", "cIn"]

needs_gre = pytest.mark.skipif(not os.path.isdir(os.path.join(GRE, "include", "Kernels")),
                               reason="GRE source tree not present")


@needs_gre
@pytest.mark.parametrize("drv", DRIVERS)
def test_gre_driver_compiles_unchanged(drv, tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    inc = tmp_path / "inc"
    inc.mkdir()
    os.symlink(os.path.join(GRE, "include", "Kernels"), inc / "Kernels")
    out = tmp_path / f"test_{drv}"
    lib = os.path.join(ROOT, "graphine_amd")
    # -fpermissive: GRE's programs initialise static const double members in
    # class (PageRank.hpp:20-21), which GRE's own build also needs
    cmd = ["g++", "-std=gnu++17", "-fpermissive", "-fopenmp", "-w",
           "-I", os.path.join(ROOT, "graphine_amd", "include"), "-I", os.path.join(ROOT, "include"),
           "-I", str(inc), os.path.join(GRE, "usr", f"test_{drv}.cpp"), "-o", str(out),
           "-L", lib, "-lgre_dev", f"-Wl,-rpath,{lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert out.exists()
