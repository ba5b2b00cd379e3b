"""GPU parity: direction-optimising BFS, 64-source bitmask BFS and batched FP64
Brandes BC through the C ABI, against the oracle.  Levels are bit-exact; BC
dependencies and path counts agree to 1e-9 relative (path counts are exact
integers in FP64)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def graphs(N, O, scale, ef, seed, sym=True):
    return N.Graph.rmat(scale, ef, seed, sym), O.rmat_graph(scale, ef, seed, sym)


@pytest.mark.parametrize("scale,seed", [(12, 2), (18, 2)])
def test_bfs_levels_bit_exact(N, O, scale, seed):
    g, og = graphs(N, O, scale, 16, seed)
    rng = np.random.default_rng(seed)
    lcc = O.cc_labels(og)
    big = np.bincount(lcc).argmax()
    cand = np.nonzero(lcc == big)[0]
    for src in list(rng.choice(cand, 4, replace=False)) + [int(np.argmax(og.outdeg))]:
        lev, edges, depth = N.bfs(g, int(src))
        ref = O.bfs_levels(og, int(src))
        assert np.array_equal(lev, ref)
        assert edges == int(og.outdeg[ref >= 0].sum())
        assert depth == ref.max()


def test_bfs_directed_and_isolated_source(N, O):
    g, og = graphs(N, O, 14, 16, 7, sym=False)
    for src in (0, int(np.nonzero(og.outdeg == 0)[0][0])):
        lev, edges, depth = N.bfs(g, src)
        assert np.array_equal(lev, O.bfs_levels(og, src))


def test_bfs_graph500_validation_large(N, O):
    """Size-independent checks at scale 20: tree-edge and |dlevel| <= 1 rules."""
    g, og = graphs(N, O, 20, 16, 2)
    src = int(np.argmax(og.outdeg))
    lev, edges, depth = N.bfs(g, src)
    srcs = np.repeat(np.arange(og.nV), og.outdeg)
    lu, lv = lev[srcs], lev[og.idx]
    both = (lu >= 0) & (lv >= 0)
    assert np.all(np.abs(lu[both] - lv[both]) <= 1)
    assert np.all((lu >= 0) == (lv >= 0))              # reached set closed under edges
    # every reached non-source vertex has a neighbour one level up
    has_parent = np.zeros(og.nV, bool)
    np.logical_or.at(has_parent, og.idx[(lv == lu + 1) & (lu >= 0)], True)
    reached = lev > 0
    assert np.all(has_parent[reached])
    assert np.array_equal(lev, O.bfs_levels(og, src))


@pytest.mark.parametrize("nsrc", [1, 17, 64])
def test_msbfs_equals_single_source(N, O, nsrc):
    g, og = graphs(N, O, 14, 16, 4)
    rng = np.random.default_rng(nsrc)
    srcs = rng.choice(og.nV, nsrc, replace=False).astype(np.uint32)
    lev, edges = N.msbfs(g, srcs)
    total = 0
    for i, s in enumerate(srcs):
        ref = O.bfs_levels(og, int(s))
        assert np.array_equal(lev[i], ref)
        total += int(og.outdeg[ref >= 0].sum())
    assert edges == total


def test_msbfs_multi_chunk_rows(N, O):
    # rows above 1024 in-edges take the chunked bottom-up path (4096 in-edges
    # per wave, OR-combined); scale 16 has rows of up to ~9,600 in-edges
    g, og = graphs(N, O, 16, 16, 4)
    assert int(og.outdeg.max()) > 2 * 4096
    rng = np.random.default_rng(16)
    srcs = rng.choice(og.nV, 64, replace=False).astype(np.uint32)
    for rep in range(2):  # the second call reuses the long-row accumulators
        lev, edges = N.msbfs(g, srcs)
        total = 0
        for i, s in enumerate(srcs):
            ref = O.bfs_levels(og, int(s))
            assert np.array_equal(lev[i], ref)
            total += int(og.outdeg[ref >= 0].sum())
        assert edges == total


@pytest.mark.parametrize("nsrc", [1, 5, 70])
def test_bc_matches_brandes(N, O, nsrc):
    g, og = graphs(N, O, 12, 16, 4)
    rng = np.random.default_rng(nsrc)
    srcs = rng.choice(og.nV, nsrc, replace=False).astype(np.uint32)
    bc, sig, dl, lv = N.bc(g, srcs, want_parts=True)
    acc = np.zeros(og.nV)
    for i, s in enumerate(srcs):
        osig, odl, olv = O.brandes_one(og, int(s))
        assert np.array_equal(lv[i], olv)
        assert np.allclose(sig[i], osig, rtol=1e-12, atol=0)
        assert np.allclose(dl[i], odl, rtol=1e-9, atol=1e-12)
        odl = odl.copy()
        odl[s] = 0.0
        acc += odl
    assert np.abs(bc - acc).sum() <= 1e-9 * np.abs(acc).sum()


def _chain(N, O, n):
    """Undirected path 0 - 1 - ... - n-1: BFS depth n-1 from an end."""
    a = np.arange(n - 1, dtype=np.uint32)
    src = np.concatenate([a, a + 1])
    dst = np.concatenate([a + 1, a])
    return N.Graph.from_edges(n, src, dst), O.graph_from_edges(n, src, dst)


def _check_bc(N, O, g, og, srcs):
    bc, sig, dl, lv = N.bc(g, srcs, want_parts=True)
    acc = np.zeros(og.nV)
    for i, s in enumerate(srcs):
        osig, odl, olv = O.brandes_one(og, int(s))
        assert np.array_equal(lv[i], olv), i
        assert np.allclose(sig[i], osig, rtol=1e-12, atol=0), i
        assert np.allclose(dl[i], odl, rtol=1e-9, atol=1e-12), i
        odl = odl.copy()
        odl[s] = 0.0
        acc += odl
    assert np.abs(bc - acc).sum() <= 1e-9 * np.abs(acc).sum()


def test_bc_deep_chain(N, O):
    # 2 x 240 steps per chunk: many replayed batches, forward and backward
    g, og = _chain(N, O, 241)
    _check_bc(N, O, g, og, np.array([0, 240, 120, 7], np.uint32))


def test_bc_depth_limit(N, O):
    # levels are kept as uint8: a depth of 250 or more is refused, not wrapped
    g, og = _chain(N, O, 260)
    with pytest.raises(N.GreError, match="250"):
        N.bc(g, np.array([0], np.uint32))
    # the same graph from its middle stays within the limit
    _check_bc(N, O, g, og, np.array([130], np.uint32))


def test_bc_repeated_sources_three_chunks(N, O):
    # 150 sources in chunks of 64, 64, 22, with repeats inside and across chunks
    g, og = graphs(N, O, 11, 8, 7)
    rng = np.random.default_rng(11)
    srcs = rng.choice(og.nV, 150, replace=True).astype(np.uint32)
    srcs[5] = srcs[3]
    srcs[140] = srcs[3]
    _check_bc(N, O, g, og, srcs)
    # a second call of another size on the same graph reuses the captured steps
    bc2, _, _, _ = N.bc(g, srcs[:70])
    bc3, _, _, _ = N.bc(g, srcs[:70])
    assert np.array_equal(bc2, bc3)


def test_msbfs_chunk_pairs_change_nothing(N, O):
    """Every long-row pair of the bottom-up pass, tiny ones included (which
    chunk most rows of this graph), gives the oracle's levels; a narrow call
    leaves the pair untuned, the first call with 32 or more sources times the
    candidates and keeps one of them."""
    g, og = graphs(N, O, 12, 16, 3)
    rng = np.random.default_rng(3)
    srcs = rng.choice(og.nV, 40, replace=False).astype(np.uint32)
    assert N.traversal_chunks(g, N.CHUNKS_MSBFS) == (0, 0)
    lev4, _ = N.msbfs(g, srcs[:4])
    assert N.traversal_chunks(g, N.CHUNKS_MSBFS) == (0, 0)
    lev0, e0 = N.msbfs(g, srcs)
    assert np.array_equal(lev0[:4], lev4)
    picked = N.traversal_chunks(g, N.CHUNKS_MSBFS)
    assert picked in {(1024, 4096), (256, 1024), (4096, 16384), (512, 2048)}
    refs = [O.bfs_levels(og, int(s)) for s in srcs]
    for i in range(len(srcs)):
        assert np.array_equal(lev0[i], refs[i])
    for pair in [(8, 16), (16, 300), (64, 64), (1024, 4096)]:
        N.set_traversal_chunks(g, N.CHUNKS_MSBFS, *pair)
        lev, e = N.msbfs(g, srcs)
        assert N.traversal_chunks(g, N.CHUNKS_MSBFS) == pair
        assert np.array_equal(lev, lev0) and e == e0, pair
    N.set_traversal_chunks(g, N.CHUNKS_MSBFS, 0)
    assert N.traversal_chunks(g, N.CHUNKS_MSBFS) == (0, 0)


def test_bc_forward_chunk_pairs_change_nothing(N, O):
    """The forward pair moves only exact integer path counts: every pair gives
    bit-identical path counts, levels, dependencies and BC."""
    g, og = graphs(N, O, 12, 16, 4)
    rng = np.random.default_rng(4)
    srcs = rng.choice(og.nV, 70, replace=False).astype(np.uint32)
    ref = N.bc(g, srcs, want_parts=True)
    assert N.traversal_chunks(g, N.CHUNKS_BC_FORWARD)[0] in (256, 512, 1024, 2048)
    for pair in [(8, 8), (32, 100), (512, 512), (2048, 2048)]:
        N.set_traversal_chunks(g, N.CHUNKS_BC_FORWARD, *pair)
        got = N.bc(g, srcs, want_parts=True)
        for a, b in zip(ref, got):
            assert np.array_equal(a, b), pair
    osig, odl, olv = O.brandes_one(og, int(srcs[0]))
    assert np.array_equal(ref[3][0], olv)
    assert np.allclose(ref[2][0], odl, rtol=1e-9, atol=1e-12)


def test_traversal_chunks_argument_checks(N):
    g = N.Graph.rmat(8, 8, 1, True)
    with pytest.raises(N.GreError):
        N.set_traversal_chunks(g, 5, 16, 16)
    with pytest.raises(N.GreError):
        N.set_traversal_chunks(g, N.CHUNKS_MSBFS, 16, 0)
    with pytest.raises(N.GreError):
        N.set_traversal_chunks(g, N.CHUNKS_BC_FORWARD, 1 << 31, 16)
