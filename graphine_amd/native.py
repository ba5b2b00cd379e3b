"""ctypes binding of libgre_dev.so (the C ABI declared in include/gre_dev.h).

This is plumbing for tests and the benchmark: it passes numpy buffers to the
C ABI and raises on any non-zero status.  There is no fallback: if the HIP
library is missing or fails to load, every entry point raises.
"""
import atexit
import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgre_dev.so")

PROG = {"PageRank": 0, "PageRank_delta": 1, "BFS": 2, "ConnectedComponents": 3, "SSSP": 4,
        "Forward": 5, "Backward": 6, "CountIn": 7}
PAGERANK, PAGERANK_DELTA, BFS, CC, SSSP, BC_FORWARD, BC_BACKWARD, COUNTIN = range(8)
DIR_OUT, DIR_IN, DIR_INOUT = range(3)
ARR_SCATTER, ARR_COMBINE, ARR_VERTEX, ARR_AUX = range(4)
F32, F64 = 0, 1

VTX_DTYPE = np.dtype([("level", np.int32), ("numPaths", np.int32)])

# dtype of (scatter, combine, vertex) arrays per program
STATE_DTYPES = {
    PAGERANK: (np.float64, np.float64, np.float64),
    PAGERANK_DELTA: (np.float64, np.float64, np.float64),
    BFS: (np.uint64, np.uint64, np.uint64),
    CC: (np.uint64, np.uint64, np.uint64),
    SSSP: (np.uint32, np.uint32, np.uint32),
    BC_FORWARD: (np.int32, np.int32, VTX_DTYPE),
    BC_BACKWARD: (np.float64, np.float64, np.float64),
    COUNTIN: (np.int32, np.int32, np.int32),
}

_lib = None

# Handles still open when the interpreter exits are closed while the HIP
# runtime is up (engines and PageRank handles before the graphs they use);
# after that, finalisers do nothing.
_live_handles = weakref.WeakSet()
_live_graphs = weakref.WeakSet()
_exiting = False


def _close_all():
    global _exiting
    for pool in (_live_handles, _live_graphs):
        for o in list(pool):
            try:
                o.close()
            except Exception:
                pass
    _exiting = True


atexit.register(_close_all)


class GreError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gre status {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GreError(1, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    vp, i32, i64, u32, u64, dbl = (C.c_void_p, C.c_int, C.c_int64, C.c_uint32, C.c_uint64,
                                   C.c_double)
    pp = C.POINTER(C.c_void_p)
    sigs = {
        "gre_last_error": ([], C.c_char_p),
        "gre_init": ([i32], i32),
        "gre_set_stream": ([vp], i32),
        "gre_get_stream": ([pp], i32),
        "gre_sync": ([], i32),
        "gre_prof_enable": ([i32], i32),
        "gre_prof_reset": ([], i32),
        "gre_prof_get": ([C.c_char_p, C.POINTER(i64), C.POINTER(dbl)], i32),
        "gre_mem_in_use": ([C.POINTER(u64)], i32),
        "gre_copy_probe": ([u64, i32, C.POINTER(dbl)], i32),
        "gre_read_probe": ([u64, i32, C.POINTER(dbl)], i32),
        "gre_gather_probe": ([u64, u64, i32, C.POINTER(dbl)], i32),
        "gre_pr_get_schedule": ([vp, C.POINTER(i32), C.POINTER(i32)], i32),
        "gre_graph_from_edges": ([pp, u32, u64, vp, vp], i32),
        "gre_graph_from_csr": ([pp, u32, u64, vp, vp], i32),
        "gre_graph_rmat": ([pp, i32, i32, u64, i32], i32),
        "gre_graph_destroy": ([vp], i32),
        "gre_graph_info": ([vp, C.POINTER(u32), C.POINTER(u64)], i32),
        "gre_graph_download": ([vp, vp, vp, vp, vp, vp, vp], i32),
        "gre_graph_set_weights_u32": ([vp, vp, u64], i32),
        "gre_graph_get_weights_u32": ([vp, vp], i32),
        "gre_engine_create": ([pp, vp, i32], i32),
        "gre_engine_destroy": ([vp], i32),
        "gre_engine_config_max_supersteps": ([vp, i64], i32),
        "gre_engine_elem_size": ([vp, i32, C.POINTER(C.c_size_t)], i32),
        "gre_engine_upload": ([vp, i32, vp], i32),
        "gre_engine_download": ([vp, i32, vp], i32),
        "gre_engine_set_level": ([vp, C.c_int32], i32),
        "gre_engine_get_level": ([vp, C.POINTER(C.c_int32)], i32),
        "gre_engine_run": ([vp, i32, vp, i64, vp, i64, C.POINTER(i64)], i32),
        "gre_engine_begin": ([vp, i32, vp, i64], i32),
        "gre_engine_step": ([vp, C.POINTER(i64)], i32),
        "gre_engine_phase_times": ([vp, vp], i32),
        "gre_engine_set_mode": ([vp, i32], i32),
        "gre_engine_pull_steps": ([vp, C.POINTER(i64)], i32),
        "gre_pr_create": ([pp, vp, i32, i32, i32], i32),
        "gre_pr_destroy": ([vp], i32),
        "gre_pr_set_edge_weights": ([vp, vp, u64], i32),
        "gre_pr_get_edge_weights": ([vp, vp, vp], i32),
        "gre_pr_set_personalization": ([vp, vp], i32),
        "gre_pr_run": ([vp, dbl, i32, dbl, C.POINTER(i32), C.POINTER(dbl)], i32),
        "gre_pr_run_noalloc": ([vp, dbl, i32, dbl, C.POINTER(i32), C.POINTER(dbl)], i32),
        "gre_pr_get_ranks": ([vp, vp], i32),
        "gre_pr_adjoint": ([vp, vp, i64, dbl, i32, dbl, C.POINTER(i32), C.POINTER(dbl)], i32),
        "gre_pr_gradients": ([vp, dbl, vp, vp, vp], i32),
        "gre_pr_get_adjoint": ([vp, vp], i32),
        "gre_pr_adjoint_cotangent": ([vp, vp, dbl, i32, dbl, C.POINTER(i32), C.POINTER(dbl)], i32),
        "gre_pr_set_schedule": ([vp, i32, i32], i32),
        "gre_pr_set_graphs": ([vp, i32], i32),
        "gre_pr_set_partition": ([vp, i32, u32], i32),
        "gre_pr_get_partition": ([vp, C.POINTER(i32), C.POINTER(u32)], i32),
        "gre_bfs": ([vp, u32, vp, C.POINTER(u64), C.POINTER(C.c_int32)], i32),
        "gre_bfs_dev": ([vp, u32, vp, C.POINTER(u64), C.POINTER(C.c_int32)], i32),
        "gre_msbfs": ([vp, vp, i32, vp, C.POINTER(u64)], i32),
        "gre_bc": ([vp, vp, i32, vp, vp, vp, vp], i32),
        "gre_traversal_set_chunks": ([vp, i32, u64, u64], i32),
        "gre_traversal_get_chunks": ([vp, i32, C.POINTER(u64), C.POINTER(u64)], i32),
        "gre_vertex_cut_ldg": ([u32, u64, i32, i32, vp, vp, vp, vp], i32),
        "gre_vertex_cut_stats": ([u32, i32, i32, u64, vp, vp, vp, vp], i32),
        "gre_vertex_cut_lvid2vid": ([u32, i32, i32, i32, u64, vp, vp, vp, vp, u64, C.POINTER(u64)], i32),
    }
    for name, (args, res) in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res


def _ck(rc):
    if rc != 0:
        raise GreError(rc, lib().gre_last_error().decode(errors="replace"))


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def init(device=0):
    _ck(lib().gre_init(device))


def sync():
    _ck(lib().gre_sync())


def stream_ptr():
    s = C.c_void_p()
    _ck(lib().gre_get_stream(C.byref(s)))
    return s.value


def prof_enable(on=True):
    _ck(lib().gre_prof_enable(1 if on else 0))


def prof_reset():
    _ck(lib().gre_prof_reset())


def prof_get(name):
    n = C.c_int64()
    ms = C.c_double()
    _ck(lib().gre_prof_get(name.encode(), C.byref(n), C.byref(ms)))
    return n.value, ms.value


def copy_probe(nbytes, reps=10):
    ms = C.c_double()
    _ck(lib().gre_copy_probe(nbytes, reps, C.byref(ms)))
    return ms.value


def read_probe(nbytes, reps=10):
    ms = C.c_double()
    _ck(lib().gre_read_probe(nbytes, reps, C.byref(ms)))
    return ms.value


def gather_probe(table_entries, gathers=1 << 28, reps=5):
    """ms per pass of `gathers` random 4-byte gathers from a table of table_entries floats."""
    ms = C.c_double()
    _ck(lib().gre_gather_probe(table_entries, gathers, reps, C.byref(ms)))
    return ms.value


def mem_in_use():
    b = C.c_uint64()
    _ck(lib().gre_mem_in_use(C.byref(b)))
    return b.value


class Graph:
    """Device-resident CSR + CSC."""

    def __init__(self, handle):
        self.h = handle
        _live_graphs.add(self)
        nv = C.c_uint32()
        ne = C.c_uint64()
        _ck(lib().gre_graph_info(self.h, C.byref(nv), C.byref(ne)))
        self.nV, self.nE = nv.value, ne.value

    @classmethod
    def rmat(cls, scale, ef=16, seed=1, sym=False, scramble=False):
        h = C.c_void_p()
        _ck(lib().gre_graph_rmat(C.byref(h), scale, ef, seed, (1 if sym else 0) | (2 if scramble else 0)))
        return cls(h)

    @classmethod
    def from_csr(cls, nV, off, idx):
        off = np.ascontiguousarray(off, np.uint64)
        idx = np.ascontiguousarray(idx, np.uint32)
        h = C.c_void_p()
        _ck(lib().gre_graph_from_csr(C.byref(h), nV, len(idx), _p(off), _p(idx)))
        return cls(h)

    @classmethod
    def from_edges(cls, nV, src, dst):
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        h = C.c_void_p()
        _ck(lib().gre_graph_from_edges(C.byref(h), nV, len(src), _p(src), _p(dst)))
        return cls(h)

    def download(self, csc=True, ceid=False):
        off = np.zeros(self.nV + 1, np.uint64)
        idx = np.zeros(self.nE, np.uint32)
        coff = np.zeros(self.nV + 1, np.uint64) if csc else None
        cidx = np.zeros(self.nE, np.uint32) if csc else None
        ce = np.zeros(self.nE, np.uint64) if ceid else None
        od = np.zeros(self.nV, np.int32)
        _ck(lib().gre_graph_download(self.h, _p(off), _p(idx), _p(coff), _p(cidx), _p(ce), _p(od)))
        return dict(off=off, idx=idx, coff=coff, cidx=cidx, ceid=ce, outdeg=od)

    def outdeg(self):
        od = np.zeros(self.nV, np.int32)
        _ck(lib().gre_graph_download(self.h, None, None, None, None, None, _p(od)))
        return od

    def set_weights_u32(self, w=None, seed=3):
        if w is not None:
            w = np.ascontiguousarray(w, np.uint32)
        _ck(lib().gre_graph_set_weights_u32(self.h, _p(w), seed))

    def weights_u32(self):
        w = np.zeros(self.nE, np.uint32)
        _ck(lib().gre_graph_get_weights_u32(self.h, _p(w)))
        return w

    def close(self):
        if self.h:
            _ck(lib().gre_graph_destroy(self.h))
            self.h = None

    def __del__(self):
        if _exiting:
            return
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """SynchronousEngine<VC> for one of GRE's vertex programs, on the device."""

    def __init__(self, graph, prog):
        self.g = graph
        self.prog = prog
        self.h = C.c_void_p()
        _ck(lib().gre_engine_create(C.byref(self.h), graph.h, prog))
        _live_handles.add(self)

    max_ss = 1024

    def config_max_supersteps(self, n):
        _ck(lib().gre_engine_config_max_supersteps(self.h, n))
        self.max_ss = n if n > 0 else 1024

    def _dtype(self, which):
        if which == ARR_AUX:
            return VTX_DTYPE
        return STATE_DTYPES[self.prog][which]

    def download(self, which):
        a = np.zeros(self.g.nV, self._dtype(which))
        _ck(lib().gre_engine_download(self.h, which, _p(a)))
        return a

    def upload(self, which, arr):
        a = np.ascontiguousarray(arr, self._dtype(which))
        _ck(lib().gre_engine_upload(self.h, which, _p(a)))

    def set_level(self, level):
        _ck(lib().gre_engine_set_level(self.h, level))

    def level(self):
        v = C.c_int32()
        _ck(lib().gre_engine_get_level(self.h, C.byref(v)))
        return v.value

    def run(self, direction=DIR_OUT, srcs=None, max_counts=4096):
        s = None if srcs is None else np.ascontiguousarray(srcs, np.uint32)
        counts = np.zeros(max_counts, np.int64)
        ss = C.c_int64()
        _ck(lib().gre_engine_run(self.h, direction, _p(s), 0 if s is None else len(s), _p(counts),
                                 max_counts, C.byref(ss)))
        # a run that stops at max_supersteps recorded ss counts, one that voted
        # to halt recorded ss + 1 (the last one is 0)
        n = ss.value if ss.value >= self.max_ss else ss.value + 1
        return ss.value, counts[:min(n, max_counts)].copy()

    AUTO, PUSH, PULL = 0, 1, 2

    def set_mode(self, mode):
        _ck(lib().gre_engine_set_mode(self.h, mode))

    def pull_steps(self):
        v = C.c_int64()
        _ck(lib().gre_engine_pull_steps(self.h, C.byref(v)))
        return v.value

    def phase_times(self):
        t = np.zeros(5, np.float64)
        _ck(lib().gre_engine_phase_times(self.h, _p(t)))
        return t

    def close(self):
        if self.h:
            _ck(lib().gre_engine_destroy(self.h))
            self.h = None

    def __del__(self):
        if _exiting:
            return
        try:
            self.close()
        except Exception:
            pass


class PageRank:
    """Forward / adjoint / gradients of (personalised, weighted) PageRank."""

    def __init__(self, graph, dtype=F32, k=1, weighted=False):
        self.g = graph
        self.dtype = dtype
        self.k = k
        self.np_dtype = np.float64 if dtype == F64 else np.float32
        self.h = C.c_void_p()
        _ck(lib().gre_pr_create(C.byref(self.h), graph.h, dtype, k, 1 if weighted else 0))
        _live_handles.add(self)

    SCHEDULES = {0: "auto", 1: "tile", 2: "tile_nt", 3: "compact", 4: "compact_nt", 5: "part",
                 6: "part_nt", 7: "rows", 8: "bin"}

    def set_schedule(self, schedule, tile_items=0):
        _ck(lib().gre_pr_set_schedule(self.h, schedule, tile_items))

    def schedule(self):
        s, t = C.c_int(), C.c_int()
        _ck(lib().gre_pr_get_schedule(self.h, C.byref(s), C.byref(t)))
        return self.SCHEDULES.get(s.value, str(s.value)), t.value

    def set_graphs(self, on):
        _ck(lib().gre_pr_set_graphs(self.h, 1 if on else 0))

    def set_partition(self, ranges, shared_sources=0):
        _ck(lib().gre_pr_set_partition(self.h, int(ranges), int(shared_sources)))

    def partition(self):
        r, h = C.c_int(), C.c_uint32()
        _ck(lib().gre_pr_get_partition(self.h, C.byref(r), C.byref(h)))
        return r.value, h.value

    def set_edge_weights(self, w=None, seed=0):
        if w is not None:
            w = np.ascontiguousarray(w, np.float32)
        _ck(lib().gre_pr_set_edge_weights(self.h, _p(w), seed))

    def edge_weights(self):
        wr = np.zeros(self.g.nE, np.float32)
        wn = np.zeros(self.g.nE, np.float32)
        _ck(lib().gre_pr_get_edge_weights(self.h, _p(wr), _p(wn)))
        return wr, wn

    def set_personalization(self, p):
        if p is None:
            _ck(lib().gre_pr_set_personalization(self.h, None))
            return
        p = np.ascontiguousarray(p, self.np_dtype)
        _ck(lib().gre_pr_set_personalization(self.h, _p(p)))

    def run(self, damping=0.85, iters=20, tol=0.0):
        it = C.c_int()
        res = C.c_double()
        _ck(lib().gre_pr_run(self.h, damping, iters, tol, C.byref(it), C.byref(res)))
        return it.value, res.value

    def ranks(self):
        shape = (self.g.nV,) if self.k == 1 else (self.g.nV, self.k)
        x = np.zeros(shape, self.np_dtype)
        _ck(lib().gre_pr_get_ranks(self.h, _p(x)))
        return x

    def adjoint(self, targets, damping=0.85, iters=1000, tol=0.0):
        t = np.ascontiguousarray(targets, np.uint32)
        it = C.c_int()
        res = C.c_double()
        _ck(lib().gre_pr_adjoint(self.h, _p(t), len(t), damping, iters, tol, C.byref(it),
                                 C.byref(res)))
        return it.value, res.value

    def adjoint_cotangent(self, cot, damping=0.85, iters=1000, tol=0.0):
        c = np.ascontiguousarray(cot, self.np_dtype)
        assert c.size == self.g.nV * self.k
        it = C.c_int()
        res = C.c_double()
        _ck(lib().gre_pr_adjoint_cotangent(self.h, _p(c), damping, iters, tol, C.byref(it),
                                           C.byref(res)))
        return it.value, res.value

    def adjoint_vector(self):
        shape = (self.g.nV,) if self.k == 1 else (self.g.nV, self.k)
        x = np.zeros(shape, self.np_dtype)
        _ck(lib().gre_pr_get_adjoint(self.h, _p(x)))
        return x

    def gradients(self, damping=0.85, want_p=True):
        gW = np.zeros(self.g.nE, np.float32)
        graw = np.zeros(self.g.nE, np.float32)
        shape = (self.g.nV,) if self.k == 1 else (self.g.nV, self.k)
        gp = np.zeros(shape, self.np_dtype) if want_p else None
        _ck(lib().gre_pr_gradients(self.h, damping, _p(gW), _p(graw), _p(gp)))
        return gW, graw, gp

    def close(self):
        if self.h:
            _ck(lib().gre_pr_destroy(self.h))
            self.h = None

    def __del__(self):
        if _exiting:
            return
        try:
            self.close()
        except Exception:
            pass


def bfs(graph, src, want_levels=True):
    lev = np.zeros(graph.nV, np.int32) if want_levels else None
    e = C.c_uint64()
    d = C.c_int32()
    if want_levels:
        _ck(lib().gre_bfs(graph.h, src, _p(lev), C.byref(e), C.byref(d)))
    else:
        _ck(lib().gre_bfs_dev(graph.h, src, None, C.byref(e), C.byref(d)))
    return lev, e.value, d.value


def msbfs(graph, srcs, want_levels=True):
    s = np.ascontiguousarray(srcs, np.uint32)
    lev = np.zeros((len(s), graph.nV), np.int32) if want_levels else None
    e = C.c_uint64()
    _ck(lib().gre_msbfs(graph.h, _p(s), len(s), _p(lev), C.byref(e)))
    return lev, e.value


CHUNKS_MSBFS, CHUNKS_BC_FORWARD = 0, 1


def set_traversal_chunks(graph, which, long_rows, chunk=0):
    """Long-row (threshold, chunk) pair of a batched traversal on this graph;
    long_rows = 0 leaves it to the first call's timing of the candidates."""
    _ck(lib().gre_traversal_set_chunks(graph.h, which, long_rows, chunk))


def traversal_chunks(graph, which):
    a, b = C.c_uint64(), C.c_uint64()
    _ck(lib().gre_traversal_get_chunks(graph.h, which, C.byref(a), C.byref(b)))
    return a.value, b.value


def bc(graph, srcs, want_parts=False):
    s = np.ascontiguousarray(srcs, np.uint32)
    out = np.zeros(graph.nV, np.float64)
    sig = np.zeros((len(s), graph.nV), np.float64) if want_parts else None
    dl = np.zeros((len(s), graph.nV), np.float64) if want_parts else None
    lv = np.zeros((len(s), graph.nV), np.int32) if want_parts else None
    _ck(lib().gre_bc(graph.h, _p(s), len(s), _p(out), _p(sig), _p(dl), _p(lv)))
    return out, sig, dl, lv


def vertex_cut_ldg(nV, nE_meta, np_, files):
    """LDG placement over np parts; files: list of (src, dst) arrays in file order.
    Returns (src, dst, part) over all files concatenated."""
    off = np.zeros(len(files) + 1, np.uint64)
    off[1:] = np.cumsum([len(f[0]) for f in files])
    src = np.ascontiguousarray(np.concatenate([f[0] for f in files]), np.uint32)
    dst = np.ascontiguousarray(np.concatenate([f[1] for f in files]), np.uint32)
    part = np.zeros(len(src), np.int32)
    _ck(lib().gre_vertex_cut_ldg(nV, nE_meta, np_, len(files), _p(off), _p(src), _p(dst), _p(part)))
    return src, dst, part


def vertex_cut_stats(nV, np_, direction, src, dst, part):
    st = np.zeros((np_, 7), np.int64)
    _ck(lib().gre_vertex_cut_stats(nV, np_, direction, len(src), _p(src), _p(dst), _p(part), _p(st)))
    return st


def vertex_cut_lvid2vid(nV, np_, direction, p, src, dst, part):
    n = C.c_uint64()
    L = lib()
    _ck(L.gre_vertex_cut_lvid2vid(nV, np_, direction, p, len(src), _p(src), _p(dst), _p(part), None, 0, C.byref(n)))
    out = np.zeros(n.value, np.uint64)
    _ck(L.gre_vertex_cut_lvid2vid(nV, np_, direction, p, len(src), _p(src), _p(dst), _p(part), _p(out), n.value,
                                  C.byref(n)))
    return out
