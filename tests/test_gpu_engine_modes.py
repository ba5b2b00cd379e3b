"""GPU parity of the engine's compute-phase forms: every shipped vertex program
run with the push form only, the pull form only and the automatic switch gives
the oracle's GRE-engine results (integers bit-exact, including the
per-superstep active counts; FP64 sums to 1e-12 relative), on R-MAT graphs and
on a graph with a hub row long enough for the block-per-row pull path.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODES = [("push", 1), ("pull", 2), ("auto", 0)]


def run_mode(N, O, g, og, prog, direction, mode, srcs=None, max_ss=0, aux=None, level=1):
    e = N.Engine(g, prog)
    e.set_mode(mode)
    if max_ss:
        e.config_max_supersteps(max_ss)
    if prog == N.BC_BACKWARD:
        e.upload(N.ARR_AUX, aux)
    e.set_level(level)
    ss, counts = e.run(direction, srcs)
    ost = O.default_state(prog, og.nV)
    oss, ocounts, _ = O.gre_run(prog, direction, og, ost, aux=aux, level=level, srcs=srcs, max_ss=max_ss)
    if mode == 2 and ss > 0:
        # every superstep with a non-empty frontier ran in the pull form (a run
        # that voted to halt ran ss + 1 of them)
        assert e.pull_steps() in (ss, ss + 1)
    if mode == 1:
        assert e.pull_steps() == 0
    return e, ss, counts, oss, ocounts, ost


def hub_graph(N, O):
    # a random graph plus 70,000 edges into vertex 3 and out of vertex 5, symmetrised
    rng = np.random.default_rng(11)
    n = 90000
    src = rng.integers(0, n, 300000)
    dst = rng.integers(0, n, 300000)
    hub_in = np.arange(10, 70010)
    src = np.concatenate([src, hub_in, np.full(70000, 5)])
    dst = np.concatenate([dst, np.full(70000, 3), hub_in])
    keep = src != dst
    s2 = np.concatenate([src[keep], dst[keep]])
    d2 = np.concatenate([dst[keep], src[keep]])
    key = np.unique(s2.astype(np.int64) * n + d2)
    s3 = (key // n).astype(np.uint32)
    d3 = (key % n).astype(np.uint32)
    return N.Graph.from_edges(n, s3, d3), O.graph_from_edges(n, s3, d3)


@pytest.mark.parametrize("mode_name,mode", MODES)
def test_modes_bfs_cc_countin(N, O, mode_name, mode):
    g, og = N.Graph.rmat(14, 16, 3, True), O.rmat_graph(14, 16, 3, True)
    src = int(np.argmax(og.outdeg))
    e, ss, counts, oss, ocounts, ost = run_mode(N, O, g, og, N.BFS, N.DIR_IN, mode, srcs=[src])
    assert ss == oss and np.array_equal(counts, ocounts)
    assert np.array_equal(e.download(N.ARR_VERTEX), ost[2])
    e, ss, counts, oss, ocounts, ost = run_mode(N, O, g, og, N.CC, N.DIR_INOUT, mode)
    assert ss == oss and np.array_equal(counts, ocounts)
    assert np.array_equal(e.download(N.ARR_VERTEX), ost[2])
    gd, ogd = N.Graph.rmat(13, 16, 4, False), O.rmat_graph(13, 16, 4, False)
    for d in (N.DIR_OUT, N.DIR_IN):
        e, ss, counts, oss, ocounts, ost = run_mode(N, O, gd, ogd, N.COUNTIN, d, mode, max_ss=1)
        assert np.array_equal(e.download(N.ARR_VERTEX), ost[2])


@pytest.mark.parametrize("mode_name,mode", MODES)
def test_modes_sssp(N, O, mode_name, mode):
    g, og = N.Graph.rmat(12, 16, 2, True), O.rmat_graph(12, 16, 2, True)
    g.set_weights_u32(None, seed=3)
    w = O.sssp_weights(og, seed=3)
    src = int(np.argmax(og.outdeg))
    for d in (N.DIR_OUT, N.DIR_IN):
        e, ss, counts, oss, ocounts, ost = run_mode(N, O, g, og, N.SSSP, d, mode, srcs=[src], aux=w)
        assert ss == oss and np.array_equal(counts, ocounts)
        assert np.array_equal(e.download(N.ARR_VERTEX), ost[1])


@pytest.mark.parametrize("mode_name,mode", MODES)
def test_modes_pagerank(N, O, mode_name, mode):
    g, og = N.Graph.rmat(13, 16, 1, False), O.rmat_graph(13, 16, 1, False)
    e, ss, counts, oss, ocounts, ost = run_mode(N, O, g, og, N.PAGERANK, N.DIR_OUT, mode, max_ss=10)
    assert ss == oss == 10 and np.array_equal(counts, ocounts)
    assert np.allclose(e.download(N.ARR_VERTEX), ost[2], rtol=1e-12, atol=0)


@pytest.mark.parametrize("mode_name,mode", MODES)
def test_modes_bc_pipeline(N, O, mode_name, mode):
    g, og = N.Graph.rmat(12, 16, 4, True), O.rmat_graph(12, 16, 4, True)
    src = int(np.argmax(og.outdeg))
    f, ss, counts, oss, ocounts, fst = run_mode(N, O, g, og, N.BC_FORWARD, N.DIR_OUT, mode, srcs=[src])
    assert ss == oss and np.array_equal(counts, ocounts)
    vtx = f.download(N.ARR_VERTEX)
    assert np.array_equal(vtx, fst[2])
    b = N.Engine(g, N.BC_BACKWARD)
    b.set_mode(mode)
    b.upload(N.ARR_AUX, vtx)
    bst = O.default_state(O.BC_BACKWARD, og.nV)
    for i in range(f.level() - 2, -1, -1):
        b.set_level(i)
        b.run(N.DIR_IN, None)
        O.gre_run(O.BC_BACKWARD, O.DIR_IN, og, bst, aux=fst[2], level=i)
    assert np.allclose(b.download(N.ARR_VERTEX), bst[2], rtol=1e-12, atol=1e-12)


def test_modes_hub_rows(N, O):
    # rows of 70,000 entries take the block-per-row pull path
    g, og = hub_graph(N, O)
    assert og.indeg.max() > 65536
    g.set_weights_u32(None, seed=5)
    w = O.sssp_weights(og, seed=5)
    for mode in (1, 2):
        e, ss, counts, oss, ocounts, ost = run_mode(N, O, g, og, N.CC, N.DIR_INOUT, mode)
        assert ss == oss and np.array_equal(counts, ocounts)
        assert np.array_equal(e.download(N.ARR_VERTEX), ost[2])
        e, ss, counts, oss, ocounts, ost = run_mode(N, O, g, og, N.BFS, N.DIR_IN, mode, srcs=[7])
        assert ss == oss and np.array_equal(counts, ocounts)
        assert np.array_equal(e.download(N.ARR_VERTEX), ost[2])
        e, ss, counts, oss, ocounts, ost = run_mode(N, O, g, og, N.SSSP, N.DIR_OUT, mode, srcs=[3], aux=w)
        assert ss == oss and np.array_equal(e.download(N.ARR_VERTEX), ost[1])
        e, ss, counts, oss, ocounts, ost = run_mode(N, O, g, og, N.PAGERANK, N.DIR_OUT, mode, max_ss=5)
        assert np.allclose(e.download(N.ARR_VERTEX), ost[2], rtol=1e-12, atol=0)


def test_partitioned_pagerank_scatter_upload_between_steps(N):
    # the partitioned pull keeps a compact copy of GRE PageRank's scatter data
    # for the whole run (the program never changes it); a scatter upload
    # between supersteps must reach the next superstep, as in the push form
    import ctypes as C
    g = N.Graph.rmat(19, 16, 1, False)
    out = {}
    for mode in (N.Engine.PULL, N.Engine.PUSH):
        e = N.Engine(g, N.PAGERANK)
        e.set_mode(mode)
        L = N.lib()
        act = C.c_int64()
        assert L.gre_engine_begin(e.h, N.DIR_OUT, None, 0) == 0
        for step in range(4):
            if step == 2:
                sc = e.download(N.ARR_SCATTER)
                e.upload(N.ARR_SCATTER, sc * 0.5)
            assert L.gre_engine_step(e.h, C.byref(act)) == 0
        out[mode] = e.download(N.ARR_VERTEX)
        if mode == N.Engine.PULL:
            assert e.pull_steps() == 4
        e.close()
    g.close()
    assert np.allclose(out[N.Engine.PULL], out[N.Engine.PUSH], rtol=1e-12, atol=0)
    # and the halving took effect: the ranks differ from an unmodified run
    e = N.Engine(g2 := N.Graph.rmat(19, 16, 1, False), N.PAGERANK)
    e.config_max_supersteps(4)
    e.run(N.DIR_OUT)
    plain = e.download(N.ARR_VERTEX)
    e.close()
    g2.close()
    assert not np.allclose(plain, out[N.Engine.PULL], rtol=1e-6, atol=0)


def test_modes_pagerank_partitioned(N, O):
    # from 2^22 edges GRE's PageRank takes the binned FP64 pull (its
    # complete-frontier supersteps: exact two-word fixed-point row sums)
    g, og = N.Graph.rmat(19, 16, 1, False), O.rmat_graph(19, 16, 1, False)
    assert len(og.idx) >= 1 << 22
    for mode in (2, 0):
        e, ss, counts, oss, ocounts, ost = run_mode(N, O, g, og, N.PAGERANK, N.DIR_OUT, mode, max_ss=10)
        assert ss == oss == 10 and np.array_equal(counts, ocounts)
        assert e.pull_steps() == 10
        assert np.allclose(e.download(N.ARR_VERTEX), ost[2], rtol=1e-12, atol=0)


def test_modes_pagerank_partitioned_long_rows(N, O):
    """GRE's PageRank program on the engine's large-graph pull (the binned FP64
    pull) where a destination has 1.2 M in-edges (73 scatter chunks of 16,384
    out-degree-1 sources each write 1,024 group sums into its bin) and a source
    has 300,000 out-edges (its out-row cut into two chunks); R-MAT edges around
    them.  Counts exact; every vertex but the hub to 1e-12 of the oracle.  The
    hub is pinned to its exact value: GRE's PageRank scatters its init value
    0.15/outdeg every superstep, so the hub holds 0.15 + 0.85 * the sum over
    its in-neighbours of 0.15/outdeg, which math.fsum rounds correctly; both
    the device's sum and the oracle's one-by-one sum of the 1.2 M terms stay
    within the n*u bound of a recursive sum.  A second engine gives the same
    bits."""
    import math
    base = O.rmat_graph(19, 16, 1, False)
    n = 1 << 21
    rsrc = np.repeat(np.arange(base.nV, dtype=np.uint32), base.outdeg)
    star_in = np.arange(1 << 19, (1 << 19) + 1_200_000, dtype=np.uint32)
    src = np.concatenate([rsrc, star_in, np.full(300_000, 5, np.uint32)])
    dst = np.concatenate([base.idx, np.full(star_in.size, 7, np.uint32),
                          np.arange(1 << 20, (1 << 20) + 300_000, dtype=np.uint32)])
    g, og = N.Graph.from_edges(n, src, dst), O.graph_from_edges(n, src, dst)
    hub = 7
    assert len(og.idx) >= 1 << 22 and int(og.indeg[hub]) > 1 << 20
    ins = src[dst == hub]
    terms = [0.15 / float(og.outdeg[u]) for u in ins]
    exact = 0.15 + 0.85 * math.fsum(terms)
    bound = len(terms) * 2.0 ** -53 * exact
    others = np.arange(n) != hub
    for mode in (2, 0):
        e, ss, counts, oss, ocounts, ost = run_mode(N, O, g, og, N.PAGERANK, N.DIR_OUT, mode, max_ss=8)
        assert ss == oss == 8 and np.array_equal(counts, ocounts)
        assert e.pull_steps() == 8
        x = e.download(N.ARR_VERTEX)
        assert np.allclose(x[others], ost[2][others], rtol=1e-12, atol=0)
        assert abs(x[hub] - exact) <= bound
        assert abs(ost[2][hub] - exact) <= bound
        e.close()
        e2 = N.Engine(g, N.PAGERANK)
        e2.set_mode(mode)
        e2.config_max_supersteps(8)
        e2.run(N.DIR_OUT)
        assert np.array_equal(e2.download(N.ARR_VERTEX), x)
        e2.close()


def test_engine_pagerank_binned_multi_slice(N, O):
    """GRE's PageRank through the engine's binned FP64 pull where both bins of
    destinations take several accumulate slices: 4,500,000 out-degree-1
    sources, source s -> s mod 16,384.  The sources are staged in compact
    order (all of out-degree 1, so in vertex order) and every chunk of 16,384
    of them reaches each destination once, so no two entries share a slot:
    each of the two 8,192-row bins holds 2.25 M slots, cut into three slices
    whose two-word sums meet in 64-bit integers in HBM.  Counts exact, ranks to
    1e-12 of the oracle, every destination within the n*u bound of its exact
    value 0.15 + 0.85 * 0.15 * indegree (math.fsum), and a second engine gives
    the same bits."""
    import math
    nd, ns = 16384, 4_500_000
    n = nd + ns
    src = np.arange(nd, n, dtype=np.uint32)
    dst = (np.arange(ns, dtype=np.uint64) % nd).astype(np.uint32)
    g, og = N.Graph.from_edges(n, src, dst), O.graph_from_edges(n, src, dst)
    assert len(og.idx) >= 1 << 22
    indeg = og.indeg[:nd].astype(np.int64)
    exact = np.array([0.15 + 0.85 * math.fsum([0.15] * int(k)) for k in np.unique(indeg)])
    ex_of = dict(zip(np.unique(indeg).tolist(), exact.tolist()))
    for mode in (2, 0):
        e, ss, counts, oss, ocounts, ost = run_mode(N, O, g, og, N.PAGERANK, N.DIR_OUT, mode, max_ss=6)
        assert ss == oss == 6 and np.array_equal(counts, ocounts)
        assert e.pull_steps() == 6
        x = e.download(N.ARR_VERTEX)
        assert np.allclose(x, ost[2], rtol=1e-12, atol=0)
        for v in (0, 1, 8191, 8192, nd - 1):
            k = int(indeg[v])
            assert abs(x[v] - ex_of[k]) <= k * 2.0 ** -53 * ex_of[k], v
        e.close()
        e2 = N.Engine(g, N.PAGERANK)
        e2.set_mode(mode)
        e2.config_max_supersteps(6)
        e2.run(N.DIR_OUT)
        assert np.array_equal(e2.download(N.ARR_VERTEX), x)
        e2.close()


@pytest.mark.parametrize("mode", [2, 0])
def test_modes_pagerank_delta_partitioned(N, O, mode):
    # PageRank_delta's first superstep has a complete frontier, so from 2^22
    # edges it takes the partitioned pull (scatter data re-permuted every
    # superstep: the program rewrites it in apply); later supersteps pull or
    # push by frontier size.  Counts exact, ranks to 1e-12.
    from test_gpu_graph_engine import delta_margin
    g, og = N.Graph.rmat(19, 16, 1, False), O.rmat_graph(19, 16, 1, False)
    assert len(og.idx) >= 1 << 22
    assert delta_margin(N, g, og.outdeg, mode) > 1e-9
    e, ss, counts, oss, ocounts, ost = run_mode(N, O, g, og, N.PAGERANK_DELTA, N.DIR_OUT, mode)
    assert ss == oss and np.array_equal(counts, ocounts)
    assert e.pull_steps() >= 1
    assert np.allclose(e.download(N.ARR_VERTEX), ost[2], rtol=1e-12, atol=0)


def threshold_graph(N, O, deg):
    # six hubs -> the same `deg` targets, every target -> one sink
    hubs = np.repeat(np.arange(6), deg)
    tg = np.tile(np.arange(6, 6 + deg), 6)
    sink = 6 + deg
    src = np.concatenate([hubs, np.arange(6, 6 + deg)]).astype(np.uint32)
    dst = np.concatenate([tg, np.full(deg, sink)]).astype(np.uint32)
    return N.Graph.from_edges(sink + 1, src, dst), O.graph_from_edges(sink + 1, src, dst)


@pytest.mark.parametrize("deg", [76499, 76501])
@pytest.mark.parametrize("mode_name,mode", MODES)
def test_modes_pagerank_delta_near_threshold(N, O, deg, mode_name, mode):
    # a target's delta is 0.85 * 6 * 0.15/deg: 1.3e-5 relative above 1e-5 at
    # 76,499 (all targets activate), as far below at 76,501 (none do)
    g, og = threshold_graph(N, O, deg)
    e, ss, counts, oss, ocounts, ost = run_mode(N, O, g, og, N.PAGERANK_DELTA, N.DIR_OUT, mode)
    assert counts.tolist() == ocounts.tolist() == ([76499, 0] if deg == 76499 else [0])
    assert ss == oss
    pr = e.download(N.ARR_VERTEX)
    assert np.allclose(pr[:-1], ost[2][:-1], rtol=1e-12, atol=0)
    # the sink's rank sums `deg` messages per superstep: one at a time (the
    # oracle, GRE) and in tiles (the pull) differ by up to n * eps relative
    assert abs(pr[-1] - ost[2][-1]) <= 2 * deg * np.finfo(np.float64).eps * ost[2][-1]


def test_pagerank_delta_knife_edge_push(N, O):
    # 76,500 targets: six messages 0.15/76500 added one after another give
    # 0.85 * sum == 1e-5 exactly, and GRE's strict `delta > err` keeps every
    # target inactive.  GRE adds messages one at a time and so does the push
    # form (equal addends: every arrival order gives the same sum).  The pull
    # form sums a row split across two tiles as two partial sums, which can
    # land a few ulps off 1e-5: a summation-order difference, not tested here.
    g, og = threshold_graph(N, O, 76500)
    e, ss, counts, oss, ocounts, ost = run_mode(N, O, g, og, N.PAGERANK_DELTA, N.DIR_OUT, 1)
    assert counts.tolist() == ocounts.tolist() == [0] and ss == oss == 0
    pr = e.download(N.ARR_VERTEX)
    assert np.allclose(pr, ost[2], rtol=1e-15, atol=0)
