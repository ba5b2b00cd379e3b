"""GPU parity: PageRank / personalised PageRank forward, adjoint and gradients
through the C ABI, against the oracle (FP64) and finite differences.

Tolerances (relative L1 unless stated): FP64 device vs FP64 oracle 1e-12;
FP32 device vs FP64 oracle 1e-6 (the stated bar for PageRank); gradients in
FP32 storage 1e-5.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l1(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).sum() / max(np.abs(b).sum(), 1e-300))


@pytest.mark.parametrize("schedule,tile_items", [(1, 4), (1, 8), (1, 16), (2, 8), (3, 4), (3, 16), (4, 8),
                                                  (5, 4), (5, 8), (5, 16), (6, 8)])
def test_pagerank_f64_matches_oracle(N, O, schedule, tile_items):
    g = N.Graph.rmat(16, 16, 1, False)
    og = O.rmat_graph(16, 16, 1, False)
    pr = N.PageRank(g, N.F64)
    pr.set_schedule(schedule, tile_items)
    assert pr.schedule() == (N.PageRank.SCHEDULES[schedule], tile_items)
    it, res = pr.run(0.85, 20)
    x = pr.ranks()
    ref = O.pagerank(og, 0.85, 20)
    assert it == 20
    assert rel_l1(x, ref) < 1e-12
    assert abs(x.sum() - 1.0) < 1e-12


@pytest.mark.parametrize("dtype,schedule", [("f64", 6), ("f32", 6), ("f64", 1)])
def test_fixed_iteration_residual_is_last_change(N, dtype, schedule):
    # fixed-count runs accumulate the L1 change on the last iteration only; the
    # returned residual is the change between the last two iterates, replayed
    # from a captured graph or eager
    g = N.Graph.rmat(15, 16, 3, False)
    dt = N.F64 if dtype == "f64" else N.F32
    for graphs in (True, False):
        pr = N.PageRank(g, dt)
        pr.set_schedule(schedule, 8)
        pr.set_graphs(graphs)
        pr.run(0.85, 19)
        x19 = pr.ranks().astype(np.float64)
        it, res = pr.run(0.85, 20)
        x20 = pr.ranks().astype(np.float64)
        assert it == 20
        ref = np.abs(x20 - x19).sum()
        assert ref > 0
        assert abs(res - ref) <= 1e-9 * ref


@pytest.mark.parametrize("ranges,shared", [(8, 0), (16, 0), (32, 0), (64, 0), (128, 0), (256, 0),
                                           (64, 1000), (128, 1 << 12), (256, 1 << 12), (8, 1 << 30)])
def test_pagerank_partitioned_ranges(N, O, ranges, shared):
    # source partitions (shared hub range + tail ranges): every layout gives the
    # oracle's ranks, and two handles with the same partition give the same bits
    g = N.Graph.rmat(17, 16, 4, False)
    og = O.rmat_graph(17, 16, 4, False)
    ref = O.pagerank(og, 0.85, 20)
    xs = []
    for _ in range(2):
        pr = N.PageRank(g, N.F64)
        pr.set_schedule(5, 8)
        pr.set_partition(ranges, shared)
        assert pr.partition() == (ranges, shared)
        pr.run(0.85, 20)
        xs.append(pr.ranks())
    assert rel_l1(xs[0], ref) < 1e-12
    assert np.array_equal(xs[0], xs[1])
    pr32 = N.PageRank(g, N.F32)
    pr32.set_schedule(6, 16)
    pr32.set_partition(ranges, shared)
    pr32.run(0.85, 20)
    assert rel_l1(pr32.ranks(), ref) < 1e-6


@pytest.mark.parametrize("dtype,schedule,ti", [("f32", 3, 8), ("f32", 5, 4), ("f64", 1, 8), ("f64", 6, 8)])
def test_pagerank_graph_replay_matches_eager(N, O, dtype, schedule, ti):
    # fixed-iteration runs replay a captured hipGraph: every replay, with eager
    # work in between and after parameter changes, gives the eager bits
    g = N.Graph.rmat(15, 16, 7, False)
    dt = N.F64 if dtype == "f64" else N.F32
    ref = {}
    eager = N.PageRank(g, dt)
    eager.set_schedule(schedule, ti)
    eager.set_graphs(False)
    for it in (20, 7):
        eager.run(0.85, it)
        ref[it] = eager.ranks()
    eager.run(0.8, 20)
    ref["d"] = eager.ranks()
    pr = N.PageRank(g, dt)
    pr.set_schedule(schedule, ti)
    other = N.PageRank(g, dt)
    for rep in range(3):
        it, res = pr.run(0.85, 20)
        assert it == 20
        assert np.array_equal(pr.ranks(), ref[20]), rep
        other.run(0.85, 3, 1e-30)          # eager launches between replays
    pr.run(0.85, 7)
    assert np.array_equal(pr.ranks(), ref[7])
    pr.run(0.8, 20)
    assert np.array_equal(pr.ranks(), ref["d"])
    pr.run(0.85, 20)
    assert np.array_equal(pr.ranks(), ref[20])


def test_pagerank_f32_within_1e6_of_f64(N, O):
    g = N.Graph.rmat(18, 16, 6, False)
    og = O.rmat_graph(18, 16, 6, False)
    pr = N.PageRank(g, N.F32)
    pr.run(0.85, 20)
    ref = O.pagerank(og, 0.85, 20)
    assert rel_l1(pr.ranks(), ref) < 1e-6


def test_pagerank_autotuned_is_deterministic(N, O):
    g = N.Graph.rmat(15, 16, 3, False)
    pr = N.PageRank(g, N.F64)
    pr.run(0.85, 20)
    a = pr.ranks()
    pr.run(0.85, 20)
    b = pr.ranks()
    assert np.array_equal(a, b)


def test_pagerank_edge_cases(N, O):
    # a graph with isolated vertices, a long hub row that spans many tiles, and empty tail rows
    n = 5000
    hub_src = np.zeros(4000, np.uint32)
    hub_dst = np.arange(1, 4001, dtype=np.uint32)
    into_hub = np.arange(1, 3001, dtype=np.uint32)
    src = np.concatenate([hub_src, into_hub, np.array([10, 11], np.uint32)])
    dst = np.concatenate([hub_dst, np.zeros(3000, np.uint32), np.array([11, 10], np.uint32)])
    g = N.Graph.from_edges(n, src, dst)
    og = O.graph_from_edges(n, src, dst)
    # the adjoint too (the transposed iteration: the hub's out-row is one long CSR row)
    tg = np.array([0, 1, 10, 4000, 4999], np.uint32)
    c = np.zeros(n)
    c[tg] = 1.0
    srcs = np.repeat(np.arange(og.nV), og.outdeg)
    wn = 1.0 / og.outdeg[srcs].astype(np.float64)
    uni = np.full(n, 1.0 / n)
    lref, _ = O.ppr_adjoint(og, wn, uni, c, 0.85, tol=0.0, maxit=20)
    for sc in (1, 3, 5):
        for ti in (4, 8, 16):
            pr = N.PageRank(g, N.F64)
            pr.set_schedule(sc, ti)
            pr.run(0.85, 20)
            assert rel_l1(pr.ranks(), O.pagerank(og, 0.85, 20)) < 1e-12
            pr.adjoint(tg, 0.85, 20)
            assert rel_l1(pr.adjoint_vector(), lref) < 1e-12


def ppr_problem(O, og, seed=5, nseeds=50):
    rng = np.random.default_rng(seed)
    p = np.zeros(og.nV)
    p[rng.choice(og.nV, nseeds, replace=False)] = 1.0 / nseeds
    return p


def test_weighted_ppr_adjoint_gradients_match_oracle(N, O):
    scale = 13
    g = N.Graph.rmat(scale, 16, 5, False)
    og = O.rmat_graph(scale, 16, 5, False)
    d = 0.85
    pr = N.PageRank(g, N.F64, 1, weighted=True)
    pr.set_edge_weights(None, seed=5)
    wraw, wnorm = pr.edge_weights()
    # device lognormal weights follow the oracle's definition
    src = np.repeat(np.arange(og.nV), og.outdeg)
    for e in range(0, og.nE, max(1, og.nE // 50)):
        assert abs(wraw[e] - O.lognormal_weight(5, int(src[e]), int(og.idx[e]))) <= 1e-6 * wraw[e]
    wr = wraw.astype(np.float64)
    rs = np.bincount(src, weights=wr, minlength=og.nV)
    # an FP64 run normalises in FP64: w / rowsum, the same division as here
    wn = np.where(rs[src] > 0, wr / rs[src], 0.0)
    assert np.allclose(wnorm, wn, rtol=1e-6)
    p = ppr_problem(O, og)
    pr.set_personalization(p)
    it, res = pr.run(d, 2000, 1e-13)
    x = pr.ranks()
    wc = wn[og.ceid.astype(np.int64)]
    xref, _ = O.ppr(og, wc, p, d, tol=1e-15, maxit=5000)
    assert rel_l1(x, xref) < 1e-11
    tg = np.random.default_rng(1).choice(og.nV, 64, replace=False).astype(np.uint32)
    pr.adjoint(tg, d, 2000, 1e-13)
    lam = pr.adjoint_vector()
    c = np.zeros(og.nV)
    c[tg] = 1.0
    lref, _ = O.ppr_adjoint(og, wn, p, c, d, tol=1e-15, maxit=5000)
    assert rel_l1(lam, lref) < 1e-11
    gW, graw, gp = pr.gradients(d)
    oW, oraw = O.ppr_edge_grad(og, wn, rs, xref, lref, d)
    assert rel_l1(gW, oW) < 1e-5
    assert rel_l1(graw, oraw) < 1e-5
    gpref = ((1 - d) + d * xref[og.outdeg == 0].sum()) * lref
    assert rel_l1(gp, gpref) < 1e-11


@pytest.mark.parametrize("schedule,tile_items", [(1, 8), (3, 8), (4, 16), (5, 4), (5, 8), (6, 8)])
def test_weighted_ppr_schedules_match_oracle(N, O, schedule, tile_items):
    # compact (3, 4) and partitioned (5, 6) schedules of the weighted forward
    # gather the ranks in compact source order and multiply by the weights in
    # CSC order (compact) or in the partition's entry order (partitioned); new
    # weights between runs reach the partition's copy
    scale = 14
    g = N.Graph.rmat(scale, 16, 5, False)
    og = O.rmat_graph(scale, 16, 5, False)
    p = ppr_problem(O, og)
    pr = N.PageRank(g, N.F64, 1, weighted=True)
    pr.set_schedule(schedule, tile_items)
    if schedule in (5, 6):
        pr.set_partition(16, 1 << 10)
    pr.set_personalization(p)
    src = np.repeat(np.arange(og.nV), og.outdeg)
    for seed in (5, 9):
        pr.set_edge_weights(None, seed=seed)
        wraw, _ = pr.edge_weights()
        pr.set_graphs(seed == 9)
        it, _ = pr.run(0.85, 40)
        assert it == 40
        assert pr.schedule() == (N.PageRank.SCHEDULES[schedule], tile_items)
        # an FP64 run normalises in FP64 (the returned normalised copy is FP32)
        wr = wraw.astype(np.float64)
        rs = np.bincount(src, weights=wr, minlength=og.nV)
        wn = np.where(rs[src] > 0, wr / rs[src], 0.0)
        xref, _ = O.ppr(og, wn[og.ceid.astype(np.int64)], p, 0.85, tol=0.0, maxit=40)
        assert rel_l1(pr.ranks(), xref) < 1e-12
    p32 = N.PageRank(g, N.F32, 1, weighted=True)
    p32.set_schedule(schedule, tile_items)
    p32.set_edge_weights(None, seed=9)
    p32.set_personalization(p)
    p32.run(0.85, 40)
    assert rel_l1(p32.ranks(), xref) < 1e-6


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("schedule", [1, 3, 6])
def test_adjoint_schedules_match_oracle(N, O, weighted, schedule):
    # the compact and partitioned schedules run the adjoint over the CSR in
    # compact target ids (vertices with in-edges by descending in-degree),
    # gathering lambda from a copy in that order; tile gathers lambda directly
    scale = 14
    g = N.Graph.rmat(scale, 16, 5, False)
    og = O.rmat_graph(scale, 16, 5, False)
    p = ppr_problem(O, og)
    src = np.repeat(np.arange(og.nV), og.outdeg)
    tg = np.random.default_rng(1).choice(og.nV, 64, replace=False).astype(np.uint32)
    c = np.zeros(og.nV)
    c[tg] = 1.0
    for dt, tol in ((N.F64, 1e-12), (N.F32, 1e-6)):
        pr = N.PageRank(g, dt, 1, weighted=weighted)
        pr.set_schedule(schedule, 8)
        if schedule == 6 and dt == N.F32:
            pr.set_partition(16, 1 << 10)
        pr.set_personalization(p)
        for seed in (5, 9):
            # new weights between adjoint runs reach the partition-ordered copy
            if weighted:
                pr.set_edge_weights(None, seed=seed)
                wraw, _ = pr.edge_weights()
                wr = wraw.astype(np.float64)
                rs = np.bincount(src, weights=wr, minlength=og.nV)
                wn = np.where(rs[src] > 0, wr / rs[src], 0.0)
            else:
                wn = 1.0 / og.outdeg[src].astype(np.float64)
            pr.run(0.85, 5)
            it, _ = pr.adjoint(tg, 0.85, 30)
            assert it == 30
            lref, _ = O.ppr_adjoint(og, wn, p, c, 0.85, tol=0.0, maxit=30)
            assert rel_l1(pr.adjoint_vector(), lref) < tol
        pr.close()


def test_ppr_f32_forward_with_f64_residual(N, O):
    scale = 14
    g = N.Graph.rmat(scale, 16, 5, False)
    og = O.rmat_graph(scale, 16, 5, False)
    pr = N.PageRank(g, N.F32, 1, weighted=True)
    pr.set_edge_weights(None, seed=5)
    _, wnorm = pr.edge_weights()
    p = ppr_problem(O, og)
    pr.set_personalization(p)
    it, res = pr.run(0.85, 1000, 1e-7)
    assert res < 1e-7 and it < 1000
    wc = wnorm.astype(np.float64)[og.ceid.astype(np.int64)]
    xref, _ = O.ppr(og, wc, p, 0.85, tol=1e-15, maxit=5000)
    assert rel_l1(pr.ranks(), xref) < 1e-6


def test_gradient_central_differences_f64(N, O):
    """d(loss)/d(raw edge weight) against central differences of an FP64 forward (h = 1e-3 relative)."""
    scale = 11
    g = N.Graph.rmat(scale, 16, 5, False)
    og = O.rmat_graph(scale, 16, 5, False)
    d = 0.85
    pr = N.PageRank(g, N.F64, 1, weighted=True)
    pr.set_edge_weights(None, seed=5)
    wraw, _ = pr.edge_weights()
    p = ppr_problem(O, og, nseeds=20)
    pr.set_personalization(p)
    pr.run(d, 5000, 1e-15)
    tg = np.random.default_rng(2).choice(og.nV, 32, replace=False).astype(np.uint32)
    pr.adjoint(tg, d, 5000, 1e-15)
    _, graw, _ = pr.gradients(d)
    rng = np.random.default_rng(3)
    edges = rng.choice(og.nE, 16, replace=False)
    fd_pr = N.PageRank(g, N.F64, 1, weighted=True)
    fd_pr.set_personalization(p)
    for e in edges:
        h = 1e-3 * float(wraw[e])
        vals = []
        for sgn in (+1, -1):
            w2 = wraw.astype(np.float64).copy()
            w2[e] += sgn * h
            fd_pr.set_edge_weights(w2.astype(np.float32))
            fd_pr.run(d, 5000, 1e-15)
            vals.append(fd_pr.ranks()[tg].sum())
        # the perturbation is applied to float32 weights: use the float32 step
        w_hi = np.float32(wraw[e] + h)
        w_lo = np.float32(wraw[e] - h)
        fd = (vals[0] - vals[1]) / (float(w_hi) - float(w_lo))
        assert abs(fd - graw[e]) <= 2e-4 * max(abs(fd), 1e-9) + 1e-9


@pytest.mark.parametrize("k,dtype,sched", [(4, "f64", 0), (32, "f32", 0), (4, "f64", 7), (32, "f32", 7),
                                           (3, "f64", 1), (17, "f32", 7)])
def test_batched_ppr_equals_single_vectors(N, O, k, dtype, sched):
    scale = 13
    g = N.Graph.rmat(scale, 16, 6, False)
    og = O.rmat_graph(scale, 16, 6, False)
    dt = N.F64 if dtype == "f64" else N.F32
    npd = np.float64 if dtype == "f64" else np.float32
    rng = np.random.default_rng(7)
    P = np.zeros((og.nV, k), npd)
    for j in range(k):
        P[rng.choice(og.nV, 100, replace=False), j] = 1.0 / 100
    bp = N.PageRank(g, dt, k, weighted=True)
    bp.set_schedule(sched)
    bp.set_edge_weights(None, seed=6)
    bp.set_personalization(P)
    bp.run(0.85, 30)
    if sched:
        assert bp.schedule()[0] == N.PageRank.SCHEDULES[sched]
    else:
        assert bp.schedule()[0] in ("tile", "rows")
    X = bp.ranks()
    tg = rng.choice(og.nV, 128, replace=False).astype(np.uint32)
    bp.adjoint(tg, 0.85, 30)
    LAM = bp.adjoint_vector()
    gW, graw, gp = bp.gradients(0.85)
    sp = N.PageRank(g, dt, 1, weighted=True)
    sp.set_edge_weights(None, seed=6)
    gW_sum = np.zeros(og.nE)
    graw_sum = np.zeros(og.nE)
    tol = 1e-12 if dtype == "f64" else 2e-6
    for j in range(k):
        sp.set_personalization(np.ascontiguousarray(P[:, j]))
        sp.run(0.85, 30)
        assert rel_l1(X[:, j], sp.ranks().astype(np.float64)) < tol
        sp.adjoint(tg, 0.85, 30)
        assert rel_l1(LAM[:, j], sp.adjoint_vector().astype(np.float64)) < tol
        a, b, c = sp.gradients(0.85)
        gW_sum += a
        graw_sum += b
        assert rel_l1(gp[:, j], c.astype(np.float64)) < tol
    assert rel_l1(gW, gW_sum) < 1e-5
    assert rel_l1(graw, graw_sum) < 1e-4


@pytest.mark.parametrize("sched", [1, 7])
def test_batched_graph_replay_matches_eager(N, O, sched):
    # fixed-iteration batched forward and adjoint runs replay captured graphs:
    # every replay gives the eager bits, also after a damping change
    g = N.Graph.rmat(14, 16, 3, False)
    k = 8
    rng = np.random.default_rng(3)
    P = np.zeros((g.nV, k), np.float32)
    for j in range(k):
        P[rng.choice(g.nV, 50, replace=False), j] = 1.0 / 50
    tg = rng.choice(g.nV, 64, replace=False).astype(np.uint32)
    out = {}
    for graphs in (False, True):
        bp = N.PageRank(g, N.F32, k, weighted=True)
        bp.set_schedule(sched)
        bp.set_graphs(graphs)
        bp.set_edge_weights(None, seed=2)
        bp.set_personalization(P)
        res = []
        for d in (0.85, 0.85, 0.8, 0.85):
            bp.run(d, 15)
            x = bp.ranks()
            bp.adjoint(tg, d, 15)
            res.append((x, bp.adjoint_vector()))
        out[graphs] = res
    for (xe, le), (xg, lg) in zip(out[False], out[True]):
        assert np.array_equal(xe, xg)
        assert np.array_equal(le, lg)
    assert np.array_equal(out[True][0][0], out[True][1][0])


def test_partitioned_rows_relabelled_between_runs(N, O):
    """The partitioned forward walks its rows in a relabelled order and leaves
    the ranks in that order until they are read: reading them twice gives the
    same vertex-order ranks, a new personalisation reaches the next run, and
    the gradients read after a partitioned run equal those after a tile run."""
    scale, d = 13, 0.85
    g = N.Graph.rmat(scale, 16, 5, False)
    og = O.rmat_graph(scale, 16, 5, False)
    part = N.PageRank(g, N.F64, 1, weighted=True)
    part.set_schedule(6, 8)
    part.set_partition(16, 1 << 10)
    tile = N.PageRank(g, N.F64, 1, weighted=True)
    tile.set_schedule(1, 8)
    for h in (part, tile):
        h.set_edge_weights(None, seed=5)
    for seed in (5, 7):
        p = ppr_problem(O, og, seed=seed)
        for h in (part, tile):
            h.set_personalization(p)
            h.set_graphs(seed == 7)
            h.run(d, 60)
        x1 = part.ranks()
        assert np.array_equal(x1, part.ranks())
        assert rel_l1(x1, tile.ranks()) < 1e-12
    tg = np.random.default_rng(2).choice(og.nV, 32, replace=False).astype(np.uint32)
    for h in (part, tile):
        h.adjoint(tg, d, 60)
        h.run(d, 60)               # a fresh forward: its ranks are read by the gradients
    gW, graw, gp = part.gradients(d)
    rW, rraw, rp = tile.gradients(d)
    assert rel_l1(gW, rW) < 1e-9 and rel_l1(graw, rraw) < 1e-9 and rel_l1(gp, rp) < 1e-9


def test_partitioned_unweighted_f32_personalisation_changes(N, O):
    scale = 14
    g = N.Graph.rmat(scale, 16, 3, False)
    og = O.rmat_graph(scale, 16, 3, False)
    pr = N.PageRank(g, N.F32)
    pr.set_schedule(6, 8)
    pr.set_partition(32, 1 << 8)
    for seed, graphs in ((3, True), (4, True), (5, False)):
        p = ppr_problem(O, og, seed=seed)
        pr.set_personalization(p)
        pr.set_graphs(graphs)
        pr.run(0.85, 30)
        src = np.repeat(np.arange(og.nV), og.outdeg)
        wc = (1.0 / og.outdeg[src].astype(np.float64))[og.ceid.astype(np.int64)]
        xref, _ = O.ppr(og, wc, p, 0.85, tol=0.0, maxit=30)
        assert rel_l1(pr.ranks(), xref) < 1e-6, seed
