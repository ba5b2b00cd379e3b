"""The C ABI library loads and exports every entry point include/gre_dev.h declares (CPU)."""
import ctypes
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gre_dev.h")
LIB = os.path.join(ROOT, "graphine_amd", "libgre_dev.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gre_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "gre_engine_run" in names and "gre_pr_run" in names and len(names) >= 40


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build libgre_dev.so first (__graft_entry__.build())"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (gre_[a-z0-9_]+)", out))
    assert set(declared()) <= exported


def test_errors_without_gpu_are_reported_not_raised():
    # No device here: calls must return a status and a message, never crash.
    lib = ctypes.CDLL(LIB)
    lib.gre_last_error.restype = ctypes.c_char_p
    h = ctypes.c_void_p()
    rc = lib.gre_graph_rmat(ctypes.byref(h), 4, 4, ctypes.c_uint64(1), 0)
    if rc != 0:
        assert lib.gre_last_error()
    else:  # a GPU is present: the call worked
        lib.gre_graph_destroy(h)
    rc = lib.gre_graph_info(None, None, None)
    assert rc == 3 and b"null" in lib.gre_last_error()
