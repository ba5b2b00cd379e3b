"""GPU parity: device graph builds and GRE's vertex programs on the device
engine, through the C ABI, against the oracle's GRE-engine restatement.

Integer programs (BFS, CC, SSSP, BC forward, CountIn) must be bit-exact,
including the per-superstep active counts GRE prints; FP64 sums (PageRank,
BC backward) are order-dependent in GRE itself and are compared to 1e-12
relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dev_graph(N, O, scale, ef, seed, sym):
    g = N.Graph.rmat(scale, ef, seed, sym)
    og = O.rmat_graph(scale, ef, seed, sym)
    return g, og


@pytest.mark.parametrize("scale,ef,seed,sym", [(8, 4, 1, False), (12, 16, 2, True), (14, 16, 5, False),
                                               (16, 16, 1, False)])
def test_rmat_device_equals_oracle(N, O, scale, ef, seed, sym):
    g, og = dev_graph(N, O, scale, ef, seed, sym)
    d = g.download(csc=True, ceid=True)
    assert g.nE == og.nE
    assert np.array_equal(d["off"], og.off) and np.array_equal(d["idx"], og.idx)
    assert np.array_equal(d["coff"], og.coff) and np.array_equal(d["cidx"], og.cidx)
    assert np.array_equal(d["ceid"], og.ceid)
    assert np.array_equal(d["outdeg"], og.outdeg)


@pytest.mark.parametrize("sym", [False, True])
def test_rmat_scrambled_device_equals_oracle(N, O, sym):
    g = N.Graph.rmat(13, 16, 4, sym, scramble=True)
    og = O.rmat_graph(13, 16, 4, sym, scramble=True)
    d = g.download(csc=True)
    assert g.nE == og.nE
    assert np.array_equal(d["off"], og.off) and np.array_equal(d["idx"], og.idx)
    assert np.array_equal(d["coff"], og.coff) and np.array_equal(d["cidx"], og.cidx)


def test_graph_from_edges_keeps_input_order(N, O):
    rng = np.random.default_rng(3)
    n = 300
    src = rng.integers(0, n, 5000).astype(np.uint32)
    dst = rng.integers(0, n, 5000).astype(np.uint32)
    g = N.Graph.from_edges(n, src, dst)
    d = g.download(csc=True, ceid=True)
    og = O.graph_from_edges(n, src, dst)
    assert np.array_equal(d["off"], og.off) and np.array_equal(d["idx"], og.idx)
    # GraphInCSR::rbuild: CSC rows in input order (stable sort of input by destination)
    order = np.argsort(dst, kind="stable")
    assert np.array_equal(d["cidx"], src[order])
    pos_of_input = np.empty(len(src), np.int64)
    pos_of_input[np.argsort(src, kind="stable")] = np.arange(len(src))
    assert np.array_equal(d["ceid"], pos_of_input[order])


def test_graph_empty_and_isolated(N, O):
    g = N.Graph.from_edges(7, np.array([], np.uint32), np.array([], np.uint32))
    assert g.nE == 0
    e = N.Engine(g, N.CC)
    ss, counts = e.run(N.DIR_INOUT)
    assert counts.tolist() == [0] and ss == 0
    assert e.download(N.ARR_VERTEX).tolist() == list(range(7))


def test_sssp_weights_device_equal_oracle(N, O):
    g, og = dev_graph(N, O, 12, 16, 2, True)
    g.set_weights_u32(None, seed=3)
    assert np.array_equal(g.weights_u32(), O.sssp_weights(og, seed=3))


def run_both(N, O, g, og, prog, direction, srcs=None, max_ss=0, aux=None, level=1, state=None):
    e = N.Engine(g, prog)
    if max_ss:
        e.config_max_supersteps(max_ss)
    ost = O.default_state(prog, og.nV) if state is None else state
    if prog == N.BC_BACKWARD:
        e.upload(N.ARR_AUX, aux)
    e.set_level(level)
    ss, counts = e.run(direction, srcs)
    oss, ocounts, olevel = O.gre_run(prog, direction, og, ost, aux=aux, level=level, srcs=srcs,
                                     max_ss=max_ss)
    return e, ss, counts, oss, ocounts, ost, olevel


@pytest.mark.parametrize("scale,seed", [(12, 2), (16, 3)])
def test_engine_bfs(N, O, scale, seed):
    g, og = dev_graph(N, O, scale, 16, seed, True)
    src = int(np.argmax(og.outdeg))
    e, ss, counts, oss, ocounts, ost, _ = run_both(N, O, g, og, N.BFS, N.DIR_IN, srcs=[src])
    assert ss == oss and np.array_equal(counts, ocounts)
    assert np.array_equal(e.download(N.ARR_VERTEX), ost[2])
    assert np.array_equal(e.download(N.ARR_COMBINE), ost[1])


def test_engine_bfs_low_degree_source(N, O):
    g, og = dev_graph(N, O, 14, 16, 4, True)
    src = int(np.nonzero(og.outdeg == 1)[0][0])
    e, ss, counts, oss, ocounts, ost, _ = run_both(N, O, g, og, N.BFS, N.DIR_IN, srcs=[src])
    assert ss == oss and np.array_equal(counts, ocounts)
    assert np.array_equal(e.download(N.ARR_VERTEX), ost[2])


@pytest.mark.parametrize("scale", [12, 16])
def test_engine_cc(N, O, scale):
    g, og = dev_graph(N, O, scale, 8, 2, True)
    e, ss, counts, oss, ocounts, ost, _ = run_both(N, O, g, og, N.CC, N.DIR_INOUT)
    assert ss == oss and np.array_equal(counts, ocounts)
    lab = e.download(N.ARR_VERTEX)
    assert np.array_equal(lab, ost[2])
    assert np.array_equal(lab.astype(np.uint32), O.cc_labels(og))


@pytest.mark.parametrize("scale", [12, 16])
def test_engine_sssp(N, O, scale):
    g, og = dev_graph(N, O, scale, 16, 2, True)
    g.set_weights_u32(None, seed=3)
    w = O.sssp_weights(og, seed=3)
    src = int(np.argmax(og.outdeg))
    e, ss, counts, oss, ocounts, ost, _ = run_both(N, O, g, og, N.SSSP, N.DIR_OUT, srcs=[src], aux=w)
    assert ss == oss and np.array_equal(counts, ocounts)
    dist = e.download(N.ARR_VERTEX)
    assert np.array_equal(dist, ost[1])
    assert np.array_equal(dist, O.dijkstra(og, w, src))


def test_engine_sssp_in_direction_uses_edge_weights(N, O):
    g, og = dev_graph(N, O, 10, 16, 6, True)
    g.set_weights_u32(None, seed=9)
    w = O.sssp_weights(og, seed=9)
    e, ss, counts, oss, ocounts, ost, _ = run_both(N, O, g, og, N.SSSP, N.DIR_IN, srcs=[0], aux=w)
    assert np.array_equal(e.download(N.ARR_VERTEX), ost[1]) and np.array_equal(counts, ocounts)


@pytest.mark.parametrize("scale,sym", [(12, False), (14, True)])
def test_engine_pagerank(N, O, scale, sym):
    g, og = dev_graph(N, O, scale, 16, 1, sym)
    e, ss, counts, oss, ocounts, ost, _ = run_both(N, O, g, og, N.PAGERANK, N.DIR_OUT, max_ss=10)
    assert ss == oss == 10 and np.array_equal(counts, ocounts)
    pr = e.download(N.ARR_VERTEX)
    assert np.allclose(pr, ost[2], rtol=1e-12, atol=0)


def delta_margin(N, g, outdeg, mode, max_ss=0):
    """Runs PageRank_delta superstep by superstep on the device and returns the
    smallest relative distance of any propagating-candidate delta (a vertex with
    out-edges that received messages) from the 1e-5 activation threshold
    (PageRank.hpp:182).  Far from it, summation order cannot flip an activation."""
    import ctypes as C
    e = N.Engine(g, N.PAGERANK_DELTA)
    e.set_mode(mode)
    L = N.lib()
    act = C.c_int64(1)
    assert L.gre_engine_begin(e.h, N.DIR_OUT, None, 0) == 0
    prev = e.download(N.ARR_VERTEX)
    margin, steps = np.inf, 0
    while act.value > 0 and (max_ss == 0 or steps < max_ss):
        assert L.gre_engine_step(e.h, C.byref(act)) == 0
        cur = e.download(N.ARR_VERTEX)
        delta = cur - prev
        prev = cur
        cand = (delta > 0) & (outdeg > 0)
        if cand.any():
            margin = min(margin, float(np.min(np.abs(delta[cand] - 1e-5)) / 1e-5))
        steps += 1
    e.close()
    return margin


def test_engine_pagerank_delta(N, O):
    # every delta stays more than 1e-9 (relative) away from the threshold, so
    # the activations, and with them the per-superstep counts, are exact
    g, og = dev_graph(N, O, 12, 16, 1, False)
    assert delta_margin(N, g, og.outdeg, N.Engine.AUTO) > 1e-9
    e, ss, counts, oss, ocounts, ost, _ = run_both(N, O, g, og, N.PAGERANK_DELTA, N.DIR_OUT)
    assert ss == oss and 5 < ss < 1024 and np.array_equal(counts, ocounts)
    pr = e.download(N.ARR_VERTEX)
    assert np.allclose(pr, ost[2], rtol=1e-12, atol=0)


def test_engine_countin(N, O):
    g, og = dev_graph(N, O, 14, 16, 2, False)
    for direction in (N.DIR_OUT, N.DIR_IN):
        e, ss, counts, oss, ocounts, ost, _ = run_both(N, O, g, og, N.COUNTIN, direction, max_ss=1)
        cn = e.download(N.ARR_VERTEX)
        assert np.array_equal(cn, ost[2])
        assert cn.sum() == og.nE


def test_engine_bc_pipeline(N, O):
    """usr/test_bc.cpp:61-110 on the device: Forward run<Out> then Backward run<In> per level."""
    g, og = dev_graph(N, O, 13, 16, 4, True)
    src = int(np.argmax(og.outdeg))
    f = N.Engine(g, N.BC_FORWARD)
    f.set_level(1)
    ss, counts = f.run(N.DIR_OUT, [src])
    fst = O.default_state(O.BC_FORWARD, og.nV)
    oss, ocounts, olevel = O.gre_run(O.BC_FORWARD, O.DIR_OUT, og, fst, srcs=[src], level=1)
    assert ss == oss and np.array_equal(counts, ocounts) and f.level() == olevel
    vtx = f.download(N.ARR_VERTEX)
    assert np.array_equal(vtx, fst[2])           # int32 path counts wrap exactly like GRE's int
    b = N.Engine(g, N.BC_BACKWARD)
    b.upload(N.ARR_AUX, vtx)
    bst = O.default_state(O.BC_BACKWARD, og.nV)
    for i in range(olevel - 2, -1, -1):
        b.set_level(i)
        b.run(N.DIR_IN, None)
        O.gre_run(O.BC_BACKWARD, O.DIR_IN, og, bst, aux=fst[2], level=i)
    dep = b.download(N.ARR_VERTEX)
    assert np.allclose(dep, bst[2], rtol=1e-12, atol=1e-12)


def test_engine_max_supersteps_truncates(N, O):
    g, og = dev_graph(N, O, 12, 16, 2, True)
    e, ss, counts, oss, ocounts, ost, _ = run_both(N, O, g, og, N.CC, N.DIR_INOUT, max_ss=2)
    assert ss == oss == 2 and np.array_equal(counts, ocounts)
    assert np.array_equal(e.download(N.ARR_VERTEX), ost[2])


def test_engine_rejects_bad_input(N, O):
    g = N.Graph.rmat(8, 4, 1, False)
    with pytest.raises(N.GreError):
        N.Engine(g, 99)
    e = N.Engine(g, N.SSSP)
    with pytest.raises(N.GreError):      # no weights set
        e.run(N.DIR_OUT, [0])
    e = N.Engine(g, N.BFS)
    with pytest.raises(N.GreError):
        e.run(N.DIR_OUT, [1 << 20])
