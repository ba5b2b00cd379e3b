"""The oracle: TEST INFRASTRUCTURE ONLY.

ctypes bindings to ``oracle/_build/libgre_oracle.so`` (built from
``oracle/gre_oracle.c`` by ``oracle/Makefile``).  Only tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import this
module, and only as the checker or the timed CPU baseline.  The product path
(``graphine_amd``) never imports it.

Each function names the GRE source it restates (paths relative to the GRE
source tree).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libgre_oracle.so")
_lib = None

PAGERANK, PAGERANK_DELTA, BFS, CC, SSSP, BC_FORWARD, BC_BACKWARD, COUNTIN = range(8)
DIR_OUT, DIR_IN, DIR_INOUT = range(3)

VTX_DTYPE = np.dtype([("level", np.int32), ("numPaths", np.int32)])  # BC.hpp:12-17


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class _Graph(C.Structure):
    _fields_ = [("nV", C.c_uint32), ("off", C.c_void_p), ("idx", C.c_void_p),
                ("coff", C.c_void_p), ("cidx", C.c_void_p), ("ceid", C.c_void_p),
                ("outdeg", C.c_void_p)]


def _declare(L):
    vp, u32, u64, i32, i64, dbl, sz = (C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32,
                                       C.c_int64, C.c_double, C.c_size_t)
    L.o_mix64.argtypes, L.o_mix64.restype = [u64], u64
    L.o_rmat_keys.argtypes, L.o_rmat_keys.restype = [C.c_int, C.c_int, u64, C.c_int, vp], sz
    L.o_rmat_raw.argtypes = [C.c_int, u64, u64, u64, vp, vp]
    L.o_rmat_scramble_many.argtypes = [C.c_int, u64, u64, vp, vp]
    L.o_sort_unique_keys.argtypes = [vp, C.POINTER(sz)]
    L.o_csr_from_keys.argtypes = [u32, sz, vp, vp, vp]
    L.o_csr_from_edges.argtypes = [u32, sz, vp, vp, vp, vp]
    L.o_csc_from_csr.argtypes = [u32, vp, vp, vp, vp, vp]
    L.o_sssp_weight.argtypes, L.o_sssp_weight.restype = [u64, u32, u32], u32
    L.o_sssp_weights_csr.argtypes = [u64, u32, vp, vp, vp]
    L.o_lognormal_weight.argtypes, L.o_lognormal_weight.restype = [u64, u32, u32], dbl
    L.o_gre_run.argtypes = [C.c_int, C.c_int, C.POINTER(_Graph), vp, vp, vp, vp, vp, vp, i64,
                            i64, vp, i64]
    L.o_gre_run.restype = i64
    L.o_bfs_levels.argtypes = [u32, vp, vp, u32, vp]
    L.o_cc_unionfind.argtypes = [u32, vp, vp, vp]
    L.o_dijkstra.argtypes = [u32, vp, vp, vp, u32, vp]
    L.o_brandes_one.argtypes = [u32, vp, vp, u32, vp, vp, vp]
    L.o_pagerank.argtypes = [u32, vp, vp, vp, dbl, C.c_int, vp]
    L.o_ppr.argtypes, L.o_ppr.restype = [u32, vp, vp, vp, vp, vp, dbl, dbl, C.c_int, vp], C.c_int
    L.o_ppr_adjoint.argtypes = [u32, vp, vp, vp, vp, vp, vp, dbl, dbl, C.c_int, vp]
    L.o_ppr_adjoint.restype = C.c_int
    L.o_ppr_edge_grad.argtypes = [u32, vp, vp, vp, vp, vp, vp, dbl, vp, vp]
    L.o_gre_pagerank_mt.argtypes = [C.c_int, u32, vp, vp, vp, dbl, C.c_int, vp]
    L.o_gre_pagerank_mt.restype = dbl
    L.o_ldg_assign.argtypes = [u32, u64, C.c_int, C.c_int, vp, vp, vp, vp]
    L.o_ldg_assign.restype = C.c_int
    L.o_vertex_cut_stats.argtypes = [u32, C.c_int, C.c_int, u64, vp, vp, vp, vp]
    L.o_gre_pagerank_np.argtypes = [C.c_int, C.c_int, vp, u32, vp, vp, vp, dbl, C.c_int, vp, vp, vp]
    L.o_gre_pagerank_np.restype = C.c_int


# --------------------------------------------------------------------------
# graphs
# --------------------------------------------------------------------------
class Graph:
    """CSR + CSC (+ CSC->CSR position map) + out-degrees, as GRE's GraphInCSR
    holds them at np=1 (include/Graph/LocalGraph.hpp:23-200)."""

    def __init__(self, nV, off, idx):
        self.nV = int(nV)
        self.off = np.ascontiguousarray(off, dtype=np.uint64)
        self.idx = np.ascontiguousarray(idx, dtype=np.uint32)
        self.nE = int(self.off[-1])
        self.coff = np.zeros(self.nV + 1, np.uint64)
        self.cidx = np.zeros(self.nE, np.uint32)
        self.ceid = np.zeros(self.nE, np.uint64)
        lib().o_csc_from_csr(self.nV, _p(self.off), _p(self.idx), _p(self.coff), _p(self.cidx),
                             _p(self.ceid))
        self.outdeg = np.diff(self.off).astype(np.int32)
        self.indeg = np.diff(self.coff).astype(np.int32)

    def cstruct(self):
        g = _Graph(self.nV, _p(self.off).value, _p(self.idx).value, _p(self.coff).value,
                   _p(self.cidx).value, _p(self.ceid).value, _p(self.outdeg).value)
        return g


def mix64(k):
    return int(lib().o_mix64(int(k) & 0xFFFFFFFFFFFFFFFF))


def rmat_raw(scale, seed, n, first=0):
    u = np.zeros(n, np.uint32)
    v = np.zeros(n, np.uint32)
    lib().o_rmat_raw(scale, seed, first, n, _p(u), _p(v))
    return u, v


def rmat_keys(scale, ef=16, seed=1, sym=False, scramble=False):
    cap = (ef << scale) * (2 if sym else 1)
    keys = np.empty(cap, np.uint64)
    n = lib().o_rmat_keys(scale, ef, seed, (1 if sym else 0) | (2 if scramble else 0), _p(keys))
    return keys[:n].copy()


def rmat_scramble(scale, seed, x):
    """The generator's label bijection of [0, 2^scale), applied to x."""
    x = np.ascontiguousarray(x, np.uint32)
    y = np.empty_like(x)
    lib().o_rmat_scramble_many(scale, seed, len(x), _p(x), _p(y))
    return y


def csr_from_keys(nV, keys):
    off = np.zeros(nV + 1, np.uint64)
    idx = np.zeros(len(keys), np.uint32)
    lib().o_csr_from_keys(nV, len(keys), _p(keys), _p(off), _p(idx))
    return off, idx


def rmat_graph(scale, ef=16, seed=1, sym=False, scramble=False):
    """Seeded R-MAT (a,b,c,d)=(.57,.19,.19,.05); self-loops and duplicates removed;
    scramble: vertex labels permuted by rmat_scramble."""
    nV = 1 << scale
    keys = rmat_keys(scale, ef, seed, sym, scramble)
    off, idx = csr_from_keys(nV, keys)
    return Graph(nV, off, idx)


def graph_from_edges(nV, src, dst):
    """GRE np=1 ingress order: rows keep input order (LocalGraph.hpp:66-107)."""
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    off = np.zeros(nV + 1, np.uint64)
    idx = np.zeros(len(src), np.uint32)
    lib().o_csr_from_edges(nV, len(src), _p(src), _p(dst), _p(off), _p(idx))
    return Graph(nV, off, idx)


def sssp_weights(g, seed=3):
    w = np.zeros(g.nE, np.uint32)
    lib().o_sssp_weights_csr(seed, g.nV, _p(g.off), _p(g.idx), _p(w))
    return w


def sssp_weight(seed, u, v):
    return int(lib().o_sssp_weight(seed, u, v))


def lognormal_weight(seed, u, v):
    return float(lib().o_lognormal_weight(seed, u, v))


# --------------------------------------------------------------------------
# GRE engine restatement
# --------------------------------------------------------------------------
def default_state(prog, nV):
    """Arrays exactly as each program's alloc() leaves them
    (include/Kernels/*.hpp alloc(), VertexCode.hpp:132-178)."""
    if prog in (PAGERANK, PAGERANK_DELTA):
        return np.zeros(nV, np.float64), np.zeros(nV, np.float64), np.zeros(nV, np.float64)
    if prog == BFS:
        m = np.uint64(0xFFFFFFFFFFFFFFFF)
        return np.zeros(nV, np.uint64), np.full(nV, m, np.uint64), np.full(nV, m, np.uint64)
    if prog == CC:
        m = np.uint64(0xFFFFFFFFFFFFFFFF)
        return np.zeros(nV, np.uint64), np.full(nV, m, np.uint64), np.full(nV, m, np.uint64)
    if prog == SSSP:
        st = np.full(nV, 0xFFFFFFFF, np.uint32)
        return np.zeros(nV, np.uint32), st, st          # states aliases C and X
    if prog == BC_FORWARD:
        x = np.zeros(nV, VTX_DTYPE)
        x["level"] = -1
        return np.zeros(nV, np.int32), np.zeros(nV, np.int32), x
    if prog == BC_BACKWARD:
        return np.zeros(nV, np.float64), np.zeros(nV, np.float64), np.zeros(nV, np.float64)
    if prog == COUNTIN:
        return np.zeros(nV, np.int32), np.zeros(nV, np.int32), np.zeros(nV, np.int32)
    raise ValueError(prog)


def gre_run(prog, direction, g, state, aux=None, level=1, srcs=None, max_ss=0):
    """SynchronousEngine<prog>::run<dir>() at np=1 (include/Engine/Engine.hpp:180-406).
    Returns (superstep_count, active_counts, level)."""
    S, Cm, X = state
    gs = g.cstruct()
    lv = C.c_int32(level)
    cap = max(int(max_ss) if max_ss > 0 else 1024, 1)
    counts = np.zeros(cap, np.int64)
    s = None if srcs is None else np.ascontiguousarray(srcs, np.uint32)
    ns = 0 if s is None else len(s)
    ss = lib().o_gre_run(prog, direction, C.byref(gs), _p(S), _p(Cm), _p(X), _p(aux),
                         C.cast(C.pointer(lv), C.c_void_p), _p(s), ns, max_ss, _p(counts), cap)
    n = min(ss + 1, cap)
    return int(ss), counts[:n].copy(), int(lv.value)


# --------------------------------------------------------------------------
# textbook references
# --------------------------------------------------------------------------
def bfs_levels(g, src):
    lv = np.zeros(g.nV, np.int32)
    lib().o_bfs_levels(g.nV, _p(g.off), _p(g.idx), int(src), _p(lv))
    return lv


def cc_labels(g):
    lab = np.zeros(g.nV, np.uint32)
    lib().o_cc_unionfind(g.nV, _p(g.off), _p(g.idx), _p(lab))
    return lab


def dijkstra(g, w, src):
    d = np.zeros(g.nV, np.uint32)
    w = np.ascontiguousarray(w, np.uint32)
    lib().o_dijkstra(g.nV, _p(g.off), _p(g.idx), _p(w), int(src), _p(d))
    return d


def brandes_one(g, src):
    sig = np.zeros(g.nV, np.float64)
    dl = np.zeros(g.nV, np.float64)
    lv = np.zeros(g.nV, np.int32)
    lib().o_brandes_one(g.nV, _p(g.off), _p(g.idx), int(src), _p(sig), _p(dl), _p(lv))
    return sig, dl, lv


def pagerank(g, damping=0.85, iters=20):
    x = np.zeros(g.nV, np.float64)
    lib().o_pagerank(g.nV, _p(g.coff), _p(g.cidx), _p(g.outdeg), damping, iters, _p(x))
    return x


def ppr(g, wc, p, damping=0.85, tol=1e-12, maxit=10000):
    x = np.zeros(g.nV, np.float64)
    wc = np.ascontiguousarray(wc, np.float64)
    p = np.ascontiguousarray(p, np.float64)
    it = lib().o_ppr(g.nV, _p(g.coff), _p(g.cidx), _p(wc), _p(g.outdeg), _p(p), damping, tol,
                     maxit, _p(x))
    return x, it


def ppr_adjoint(g, wr, p, c, damping=0.85, tol=1e-12, maxit=10000):
    lam = np.zeros(g.nV, np.float64)
    wr = np.ascontiguousarray(wr, np.float64)
    p = np.ascontiguousarray(p, np.float64)
    c = np.ascontiguousarray(c, np.float64)
    it = lib().o_ppr_adjoint(g.nV, _p(g.off), _p(g.idx), _p(wr), _p(g.outdeg), _p(p), _p(c),
                             damping, tol, maxit, _p(lam))
    return lam, it


def ppr_edge_grad(g, wr, rowsum, x, lam, damping=0.85):
    gW = np.zeros(g.nE, np.float64)
    graw = np.zeros(g.nE, np.float64)
    args = [np.ascontiguousarray(a, np.float64) for a in (wr, rowsum, x, lam)]
    lib().o_ppr_edge_grad(g.nV, _p(g.off), _p(g.idx), *[_p(a) for a in args], damping,
                          _p(gW), _p(graw))
    return gW, graw


def gre_pagerank_mt(g, nthreads=12, damping=0.85, iters=20):
    """GRE-style push superstep with striped locks; returns (x, seconds)."""
    x = np.zeros(g.nV, np.float64)
    t = lib().o_gre_pagerank_mt(nthreads, g.nV, _p(g.off), _p(g.idx), _p(g.outdeg), damping,
                                iters, _p(x))
    return x, t


def gre_pagerank_np(g, np_=4, workers=12, damping=0.85, iters=20):
    """GRE's PageRank superstep at np_ processes on one host, over the LDG
    vertex cut of the CSR rows read as np_ edge files (GraphIngress.hpp:511-600,
    1126-1171; Engine.hpp:427-700).  workers: threads per process, or a list
    (one ingress, one timed run per entry).  Returns (x, seconds, info) with
    seconds a float or a list, and info: the last run's phase seconds,
    ingress seconds, scatter and combiner messages per superstep, edges per
    part."""
    ws = [int(w) for w in (workers if isinstance(workers, (list, tuple)) else [workers])]
    x = np.zeros(g.nV, np.float64)
    secs = np.zeros(len(ws), np.float64)
    wa = np.array(ws, np.int32)
    info = np.zeros(8 + np_, np.float64)
    rc = lib().o_gre_pagerank_np(np_, len(ws), _p(wa), g.nV, _p(g.off), _p(g.idx), _p(g.outdeg), damping,
                                 iters, _p(x), _p(secs), _p(info))
    if rc != 0:
        raise ValueError("np out of range, a worker count below 1, or fewer than np*np edges")
    phases = {k: float(v) for k, v in
              zip(("refresh", "scatter_sync", "compute", "combiner_sync", "apply"), info[:5])}
    t = [float(v) for v in secs] if isinstance(workers, (list, tuple)) else float(secs[0])
    return x, t, {"phases": phases, "ingress_s": float(info[5]), "scatter_msgs": int(info[6]),
                  "combiner_msgs": int(info[7]), "part_edges": info[8:].astype(np.int64)}


def lognormal_csr(g, seed=5):
    """Raw lognormal(0,1) weights in CSR order, their row sums, and the
    row-normalised weights (FP64)."""
    L = lib()
    w = np.empty(g.nE, np.float64)
    src = np.repeat(np.arange(g.nV, dtype=np.uint32), g.outdeg)
    for e in range(g.nE):
        w[e] = L.o_lognormal_weight(seed, int(src[e]), int(g.idx[e]))
    return w


# --------------------------------------------------------------------------
# vertex-cut ingress (GraphIngress.hpp)
# --------------------------------------------------------------------------
STAT_NAMES = ("masters", "scatters", "combiners", "bi_agents", "agents", "edges", "aligned_masters")


def ldg_assign(nV, nE_meta, np_, files):
    """LDG part of every edge: files is a list of (src, dst) arrays in file
    order; reader r takes files r, r+np, ... (GraphIngress.hpp:511-600, 1126-1171)."""
    off = np.zeros(len(files) + 1, np.uint64)
    off[1:] = np.cumsum([len(f[0]) for f in files])
    src = np.ascontiguousarray(np.concatenate([f[0] for f in files]) if files else np.zeros(0), np.uint32)
    dst = np.ascontiguousarray(np.concatenate([f[1] for f in files]) if files else np.zeros(0), np.uint32)
    part = np.zeros(len(src), np.int32)
    rc = lib().o_ldg_assign(nV, nE_meta, np_, len(files), _p(off), _p(src), _p(dst), _p(part))
    if rc != 0:
        raise ValueError("capacity 0: fewer than np*np edges")
    return src, dst, part


def vertex_cut_stats(nV, np_, direction, src, dst, part):
    st = np.zeros((np_, 7), np.int64)
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    part = np.ascontiguousarray(part, np.int32)
    lib().o_vertex_cut_stats(nV, np_, direction, len(src), _p(src), _p(dst), _p(part), _p(st))
    return st
