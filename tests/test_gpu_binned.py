"""GPU parity of the binned PageRank pull (GRE_SCHED_BIN, FP32, single vector)
through the C ABI, against the FP64 oracle.

The binned pull sums the entries of one chunk to one destination (at most 16
consecutive ones) in FP64 and rounds the sum once to FP32, then sums each
row's partial sums as 64-bit fixed point (exact, then one rounding to FP32),
so it is deterministic: handles, captured replays and eager runs give the same
bits.  Tolerance against the FP64 oracle: 1e-6 relative L1 (the bar for FP32
PageRank).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BIN = 8


def rel_l1(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).sum() / max(np.abs(b).sum(), 1e-300))


def binned(N, g):
    pr = N.PageRank(g, N.F32)
    pr.set_schedule(BIN, 8)
    assert pr.schedule()[0] == "bin"
    return pr


@pytest.mark.parametrize("scale,seed", [(10, 1), (14, 2), (17, 3), (19, 6)])
def test_binned_matches_oracle(N, O, scale, seed):
    g = N.Graph.rmat(scale, 16, seed, False)
    og = O.rmat_graph(scale, 16, seed, False)
    pr = binned(N, g)
    it, res = pr.run(0.85, 20)
    x = pr.ranks()
    ref = O.pagerank(og, 0.85, 20)
    assert it == 20
    assert rel_l1(x, ref) < 1e-6
    assert abs(float(x.astype(np.float64).sum()) - 1.0) < 1e-5
    # the same bits from a second handle, from an eager run and from a replay
    pr2 = binned(N, g)
    pr2.set_graphs(False)
    pr2.run(0.85, 20)
    assert np.array_equal(pr2.ranks(), x)
    pr.run(0.85, 20)
    assert np.array_equal(pr.ranks(), x)


def test_binned_scrambled_labels(N, O):
    g = N.Graph.rmat(15, 16, 7, False, scramble=True)
    og = O.rmat_graph(15, 16, 7, False, scramble=True)
    pr = binned(N, g)
    pr.run(0.85, 20)
    assert rel_l1(pr.ranks(), O.pagerank(og, 0.85, 20)) < 1e-6


def test_binned_edge_cases(N, O):
    """A source whose out-row is cut into three scatter chunks (300,000 entries),
    a destination with 1.2 M in-edges (runs of 16,384 entries to one row, each
    cut into 1,024 groups of 16), isolated vertices,
    vertices with only in- or only out-edges, and a 2-cycle."""
    n_in, n_out = 1_200_000, 300_000
    n = n_in + 10
    src = np.concatenate([np.zeros(n_out, np.uint32),                      # 0 -> 2 .. 300,001
                          np.arange(2, n_in + 2, dtype=np.uint32),           # 2 .. 1,200,001 -> 1
                          np.array([n - 3, n - 2], np.uint32)])              # a 2-cycle
    dst = np.concatenate([np.arange(2, n_out + 2, dtype=np.uint32),
                          np.ones(n_in, np.uint32),
                          np.array([n - 2, n - 3], np.uint32)])
    g = N.Graph.from_edges(n, src, dst)
    og = O.graph_from_edges(n, src, dst)
    ref = O.pagerank(og, 0.85, 20)
    for graphs in (True, False):
        pr = binned(N, g)
        pr.set_graphs(graphs)
        it, _ = pr.run(0.85, 20)
        assert it == 20
        x = pr.ranks()
        assert rel_l1(x, ref) < 1e-6
        # vertex-level: the hub destination, a leaf of the out-star, an isolated vertex
        for v in (0, 1, 2, n_out + 1, n - 1, n - 2):
            assert abs(float(x[v]) - ref[v]) <= 1e-6 * ref[v] + 1e-12, v


def test_binned_multi_slice_bin(N, O):
    """One bin of 16,384 destinations fed by 2,000,000 out-degree-1 sources,
    source s -> s mod 16,384: the sources are staged in vertex order, so every
    chunk of 16,384 of them reaches each destination once and no entries
    share a slot; the bin's 2,000,000 slots take two accumulate slices, whose
    sums meet in 64-bit integers in HBM."""
    nd, ns = 16384, 2_000_000
    n = nd + ns
    src = np.arange(nd, n, dtype=np.uint32)
    dst = (np.arange(ns, dtype=np.uint64) % nd).astype(np.uint32)
    g = N.Graph.from_edges(n, src, dst)
    og = O.graph_from_edges(n, src, dst)
    ref = O.pagerank(og, 0.85, 20)
    for graphs in (True, False):
        pr = binned(N, g)
        pr.set_graphs(graphs)
        it, _ = pr.run(0.85, 20)
        assert it == 20
        x = pr.ranks()
        assert rel_l1(x, ref) < 1e-6
        for v in (0, 1, 5000, nd - 1, nd, n - 1):
            assert abs(float(x[v]) - ref[v]) <= 1e-6 * ref[v] + 1e-12, v


def test_binned_residual_is_last_change(N):
    g = N.Graph.rmat(15, 16, 3, False)
    for graphs in (True, False):
        pr = binned(N, g)
        pr.set_graphs(graphs)
        pr.run(0.85, 19)
        x19 = pr.ranks().astype(np.float64)
        it, res = pr.run(0.85, 20)
        x20 = pr.ranks().astype(np.float64)
        assert it == 20
        ref = np.abs(x20 - x19).sum()
        assert ref > 0
        assert abs(res - ref) <= 1e-9 * ref


def test_binned_converged_run_and_personalisation(N, O):
    scale = 14
    g = N.Graph.rmat(scale, 16, 5, False)
    og = O.rmat_graph(scale, 16, 5, False)
    src = np.repeat(np.arange(og.nV), og.outdeg)
    wc = (1.0 / og.outdeg[src].astype(np.float64))[og.ceid.astype(np.int64)]
    rng = np.random.default_rng(5)
    pr = binned(N, g)
    for seed, graphs in ((1, True), (2, True), (3, False)):
        p = np.zeros(og.nV)
        p[np.random.default_rng(seed).choice(og.nV, 100, replace=False)] = 1.0 / 100
        pr.set_personalization(p)
        pr.set_graphs(graphs)
        pr.run(0.85, 30)
        xref, _ = O.ppr(og, wc, p, 0.85, tol=0.0, maxit=30)
        assert rel_l1(pr.ranks(), xref) < 1e-6, seed
    # to a tolerance: the FP64 residual stops the run; the ranks are the converged ones
    pr.set_personalization(None)
    it, res = pr.run(0.85, 1000, 1e-6)
    assert 1 < it < 1000 and res < 1e-6
    uni = np.full(og.nV, 1.0 / og.nV)
    xref, _ = O.ppr(og, wc, uni, 0.85, tol=1e-13, maxit=10000)
    assert rel_l1(pr.ranks(), xref) < 1e-5
    del rng


def test_binned_refuses_unbounded_personalisation(N):
    g = N.Graph.rmat(12, 16, 2, False)
    pr = binned(N, g)
    p = np.full(g.nV, 2.0 / g.nV)          # L1 norm 2: its fixed-point sums are not bounded
    pr.set_personalization(p)
    with pytest.raises(N.GreError):
        pr.run(0.85, 5)
    pr.set_personalization(None)
    pr.run(0.85, 5)
    # FP64 handles refuse the schedule at their first run
    for w in (False, True):
        h = N.PageRank(g, N.F64, 1, weighted=w)
        h.set_schedule(BIN, 8)
        with pytest.raises(N.GreError):
            h.run(0.85, 5)


def test_binned_noalloc_equals_run(N):
    g = N.Graph.rmat(16, 16, 9, False)
    pr = binned(N, g)
    it0, r0 = pr.run(0.85, 20)
    x0 = pr.ranks()
    it, res = C.c_int(-1), C.c_double(-1.0)
    N.sync()
    before = N.mem_in_use()
    rc = N.lib().gre_pr_run_noalloc(pr.h, 0.85, 20, 0.0, C.byref(it), C.byref(res))
    N.sync()
    assert rc == 0, N.lib().gre_last_error()
    assert N.mem_in_use() == before
    assert it.value == it0 == 20 and res.value == r0
    assert np.array_equal(pr.ranks(), x0)


def test_binned_then_adjoint_and_gradients(N, O):
    """The adjoint of a binned handle runs the binned pull over the reversed
    graph and agrees with the tile schedule's; the gradients read the forward's
    ranks (kept in the binned row order until read) correctly."""
    scale, d = 13, 0.85
    g = N.Graph.rmat(scale, 16, 4, False)
    bh = binned(N, g)
    th = N.PageRank(g, N.F32)
    th.set_schedule(1, 8)
    tg = np.random.default_rng(3).choice(g.nV, 32, replace=False).astype(np.uint32)
    for h in (bh, th):
        h.run(d, 40)
        h.adjoint(tg, d, 40)
        h.run(d, 40)
    assert rel_l1(bh.adjoint_vector(), th.adjoint_vector().astype(np.float64)) < 1e-6
    gW, graw, gp = bh.gradients(d)
    rW, rraw, rp = th.gradients(d)
    assert rel_l1(gW, rW.astype(np.float64)) < 1e-5
    assert rel_l1(gp, rp.astype(np.float64)) < 1e-5


def test_binned_and_part_handles_share_a_graph(N, O):
    """Binned and partitioned handles on one graph, run alternately: each keeps
    its own layout and the oracle's ranks."""
    g = N.Graph.rmat(15, 16, 8, False)
    og = O.rmat_graph(15, 16, 8, False)
    ref = O.pagerank(og, 0.85, 20)
    b = binned(N, g)
    p = N.PageRank(g, N.F32)
    p.set_schedule(6, 8)
    p.set_partition(32, 1 << 8)
    for _ in range(2):
        for h in (b, p):
            h.run(0.85, 20)
            assert rel_l1(h.ranks(), ref) < 1e-6


def weighted_ref(O, og, wraw, p, iters):
    src = np.repeat(np.arange(og.nV), og.outdeg)
    wr = wraw.astype(np.float64)
    rs = np.bincount(src, weights=wr, minlength=og.nV)
    wn = np.where(rs[src] > 0, wr / rs[src], 0.0)
    xref, _ = O.ppr(og, wn[og.ceid.astype(np.int64)], p, 0.85, tol=0.0, maxit=iters)
    return xref


def test_binned_weighted_ppr_matches_oracle(N, O):
    """Weighted PPR on the binned pull: the scatter multiplies each value by its
    edge's normalised weight, read from a copy in the layout's word order; new
    weights between runs reach that copy; a tile handle gives the same ranks."""
    scale = 14
    g = N.Graph.rmat(scale, 16, 5, False)
    og = O.rmat_graph(scale, 16, 5, False)
    rng = np.random.default_rng(5)
    p = np.zeros(og.nV)
    p[rng.choice(og.nV, 50, replace=False)] = 1.0 / 50
    pr = binned_w = N.PageRank(g, N.F32, 1, weighted=True)
    pr.set_schedule(BIN, 8)
    pr.set_personalization(p)
    tile = N.PageRank(g, N.F32, 1, weighted=True)
    tile.set_schedule(1, 8)
    tile.set_personalization(p)
    for seed in (5, 9):
        for h in (binned_w, tile):
            h.set_edge_weights(None, seed=seed)
            h.set_graphs(seed == 9)
            it, _ = h.run(0.85, 40)
            assert it == 40
        wraw, _ = pr.edge_weights()
        xref = weighted_ref(O, og, wraw, p, 40)
        assert rel_l1(pr.ranks(), xref) < 1e-6, seed
        assert rel_l1(pr.ranks(), tile.ranks().astype(np.float64)) < 1e-6, seed
        assert pr.schedule()[0] == "bin"
    # user weights (a CSR-order array), a replay and an eager run give the same bits
    w = rng.lognormal(0.0, 1.0, og.nE).astype(np.float32)
    pr.set_edge_weights(w)
    pr.set_graphs(True)
    pr.run(0.85, 30)
    a = pr.ranks()
    pr.set_graphs(False)
    pr.run(0.85, 30)
    assert np.array_equal(a, pr.ranks())
    assert rel_l1(a, weighted_ref(O, og, w, p, 30)) < 1e-6


def test_binned_weighted_adjoint_gradients(N, O):
    """After a weighted binned forward the adjoint (binned pull over the
    reversed graph, weights in its word order) and the gradients read the
    forward's ranks in vertex order."""
    scale, d = 13, 0.85
    g = N.Graph.rmat(scale, 16, 5, False)
    og = O.rmat_graph(scale, 16, 5, False)
    rng = np.random.default_rng(2)
    p = np.zeros(og.nV)
    p[rng.choice(og.nV, 50, replace=False)] = 1.0 / 50
    hs = []
    for sched in (BIN, 1):
        h = N.PageRank(g, N.F32, 1, weighted=True)
        h.set_schedule(sched, 8)
        h.set_edge_weights(None, seed=5)
        h.set_personalization(p)
        h.run(d, 60)
        tg = rng.choice(og.nV, 32, replace=False).astype(np.uint32) if not hs else tg
        h.adjoint(tg, d, 60)
        h.run(d, 60)
        hs.append(h)
    assert rel_l1(hs[0].adjoint_vector(), hs[1].adjoint_vector().astype(np.float64)) < 1e-6
    gW, graw, gp = hs[0].gradients(d)
    rW, rraw, rp = hs[1].gradients(d)
    assert rel_l1(gW, rW.astype(np.float64)) < 1e-5
    assert rel_l1(graw, rraw.astype(np.float64)) < 1e-5
    assert rel_l1(gp, rp.astype(np.float64)) < 1e-5


def test_binned_bin_array_past_2_30_slots(N):
    """A bin array of more than 2^30 slots (directed R-MAT scale 26, edge factor
    24: about 1.6 G entries, 1.3 G slots): the scatter stores through four buffer resources
    of 2^30 slots each (a resource's byte offsets are 32-bit).  No oracle runs
    at this size; the ranks are checked against the partitioned schedule's on
    the same graph (both FP32, different summation orders: 1e-6 relative L1)
    and against the invariants of the iteration (sum 1, every rank at least
    (1 - d) / n)."""
    g = N.Graph.rmat(26, 24, 6, False)
    assert g.nE > 1_500_000_000  # over 2^30 slots even if a third of the entries shared one
    pr = binned(N, g)
    it, _ = pr.run(0.85, 10)
    assert it == 10
    x = pr.ranks().astype(np.float64)
    pr.close()
    part = N.PageRank(g, N.F32)
    part.set_schedule(5, 8)
    part.run(0.85, 10)
    y = part.ranks().astype(np.float64)
    part.close()
    g.close()
    n = x.size
    assert abs(x.sum() - 1.0) < 1e-4
    assert x.min() >= 0.15 / n * (1 - 1e-6)
    assert rel_l1(x, y) < 1e-6
