"""Vertex-cut placement for np processes (CPU): GRE's LDG edge assignment and
the per-part agent counts and local id layout of its multi-process ingress
(include/Graph/GraphIngress.hpp:111-247, 511-723, 1126-1171), computed on the
host by the library (gre_vertex_cut_*) and by the oracle.

Pins: two hand-worked examples (every score written out below), one of them
chosen so that GRE's operator precedence in the score (a part already holding
the source scores bal * 1.0, whatever the destination) decides the edge;
np = 1 (everything on one part, no agents); invariants of the counts on R-MAT;
and the library equal to the oracle on R-MAT graphs read as GRE's c1 harness
writes them.
"""
import numpy as np
import pytest

from graphine_amd import native as N

OUT, IN, INOUT = 0, 1, 2


def test_ldg_hand_worked(O):
    # np = 2, nE = 8 -> capacity 8/2/2 = 2.  Reader 0 (file 0), fresh state:
    #  (0,1): part0 bal 1 x (0 home 0.5 + 1 not home 0) = .5; part1 1 x (0 + 1 home .5) = .5;
    #         tie, (0+1)%2 = 1 -> part 1
    #  (0,3): part0 1 x .5 = .5; part1 bal |1-1/2| = .5, source 0 already on 1 -> .5 x 1.0 = .5;
    #         tie, (0+3)%2 = 1 -> part 1
    #  (2,1): part0 1 x (2 home .5 + 1 not on 0, not home 0) = .5; part1 bal |1-2/2| = 0 -> 0; part 0
    #  (4,5): part0 bal .5 x (.5 + 0) = .25; part1 bal 0 -> 0; part 0
    # Reader 1 (file 1), fresh state:
    #  (1,0): part0 1 x (0 + 0 home .5) = .5; part1 1 x (1 home .5 + 0) = .5; (1+0)%2 = 1 -> part 1
    #  (3,2): part0 1 x (0 + 2 home .5) = .5; part1 bal .5 x (3 home .5 + 0) = .25 -> part 0
    files = [(np.array([0, 0, 2, 4]), np.array([1, 3, 1, 5])), (np.array([1, 3]), np.array([0, 2]))]
    for impl in (N.vertex_cut_ldg, O.ldg_assign):
        src, dst, part = impl(6, 8, 2, files)
        assert part.tolist() == [1, 1, 0, 0, 1, 0]
    # part 0: edges (2,1) (4,5) (3,2): scatters {3}, combiners {1, 5}
    # part 1: edges (0,1) (0,3) (1,0): scatters {0}, combiners {0} (one bi-agent)
    expect_out = [[3, 1, 2, 0, 3, 3, 64], [3, 1, 1, 1, 1, 3, 64]]
    expect_in = [[3, 3, 3, 0, 3, 3, 64], [3, 1, 1, 1, 1, 3, 64]]
    for st in (N.vertex_cut_stats(6, 2, OUT, src, dst, part), O.vertex_cut_stats(6, 2, OUT, src, dst, part)):
        assert st.tolist() == expect_out
    assert N.vertex_cut_stats(6, 2, IN, src, dst, part).tolist() == expect_in
    assert O.vertex_cut_stats(6, 2, INOUT, src, dst, part).tolist() == expect_in
    pad = [0xFFFFFFFFFFFFFFFF] * 61
    assert N.vertex_cut_lvid2vid(6, 2, OUT, 0, src, dst, part).tolist() == [0, 2, 4] + pad + [1, 5]
    assert N.vertex_cut_lvid2vid(6, 2, IN, 0, src, dst, part).tolist() == [0, 2, 4] + pad + [1, 3, 5]
    assert N.vertex_cut_lvid2vid(6, 2, OUT, 1, src, dst, part).tolist() == [1, 3, 5] + pad + [0]
    # what GRE's PageRank prints (PageRank.hpp:117-122): 3 combiners, 2 scatters, 5 agents,
    # 6 edges, cut 5/6


def test_ldg_source_term_hides_destination(O):
    # np = 2, nE = 32 -> capacity 8, one reader.  (0,2) -> part 0 (1 x 1.0 vs 0);
    # (1,2): part0 .875 x (0 + 2 already on 0: 1.0) = .875, part1 1 x (1 home .5) = .5 -> part 0;
    # three even-even edges -> part 0 (part 1 scores 0), five edges on part 0: bal0 = .375.
    # (1,2) again: source 1 is on part 0, so GRE scores part0 .375 x 1.0 = .375 (a sum of the
    # source and destination terms would give .375 x 2.0 = .75), part1 1 x .5 = .5 -> part 1.
    files = [(np.array([0, 1, 4, 8, 12, 1]), np.array([2, 2, 6, 10, 14, 2]))]
    for impl in (N.vertex_cut_ldg, O.ldg_assign):
        _, _, part = impl(16, 32, 2, files)
        assert part.tolist() == [0, 0, 0, 0, 0, 1]


def test_ldg_one_part_is_the_one_process_graph(O):
    g = O.rmat_graph(10, 8, seed=1)
    src = np.repeat(np.arange(g.nV, dtype=np.uint32), g.outdeg)
    s, d, part = N.vertex_cut_ldg(g.nV, g.nE, 1, [(src, g.idx)])
    assert np.all(part == 0)
    st = N.vertex_cut_stats(g.nV, 1, OUT, s, d, part)
    assert st.tolist() == [[g.nV, 0, 0, 0, 0, g.nE, g.nV]]
    assert N.vertex_cut_lvid2vid(g.nV, 1, INOUT, 0, s, d, part).tolist() == list(range(g.nV))


def c1_files(O, scale, np_, seed=1, nfiles=None):
    """GRE's np-process harness layout: edge file i holds the edges whose source
    is owned by process i % np, in CSR order."""
    g = O.rmat_graph(scale, 16, seed=seed)
    src = np.repeat(np.arange(g.nV, dtype=np.uint32), g.outdeg)
    nfiles = nfiles or np_
    files = []
    for i in range(nfiles):
        sel = (src % np_ == i % np_) & ((src // np_) % (nfiles // np_) == i // np_)
        files.append((src[sel], g.idx[sel]))
    return g, files


@pytest.mark.parametrize("np_,nfiles", [(2, 2), (4, 4), (4, 8), (3, 6)])
def test_ldg_library_equals_oracle_and_counts(O, np_, nfiles):
    g, files = c1_files(O, 12, np_, nfiles=nfiles)
    s1, d1, p1 = N.vertex_cut_ldg(g.nV, g.nE, np_, files)
    s2, d2, p2 = O.ldg_assign(g.nV, g.nE, np_, files)
    assert np.array_equal(p1, p2) and np.array_equal(s1, s2)
    for direction in (OUT, IN, INOUT):
        st = N.vertex_cut_stats(g.nV, np_, direction, s1, d1, p1)
        assert np.array_equal(st, O.vertex_cut_stats(g.nV, np_, direction, s1, d1, p1))
        assert st[:, 5].sum() == g.nE and st[:, 0].sum() == g.nV
        # a scatter agent of part p is a source of one of p's edges owned elsewhere
        for p in range(np_):
            mine = p1 == p
            assert st[p, 3] <= min(st[p, 1], st[p, 2]) or direction != OUT
            lv = N.vertex_cut_lvid2vid(g.nV, np_, direction, p, s1, d1, p1)
            nm = int(st[p, 0])
            assert lv[:nm].tolist() == list(range(p, g.nV, np_))
            assert np.all(lv[nm:int(st[p, 6])] == np.uint64(0xFFFFFFFFFFFFFFFF))
            agents = lv[int(st[p, 6]):]
            assert np.all(np.diff(agents.astype(np.int64)) > 0) and np.all(agents % np_ != p)
            if direction == OUT:
                ref = np.unique(d1[mine][d1[mine] % np_ != p])
                assert np.array_equal(agents, ref.astype(np.uint64)) and len(agents) == st[p, 2]
            else:
                rs = s1[mine][s1[mine] % np_ != p]
                rd = d1[mine][d1[mine] % np_ != p]
                assert np.array_equal(agents, np.union1d(rs, rd).astype(np.uint64))
    # the same input gives the same placement
    assert np.array_equal(N.vertex_cut_ldg(g.nV, g.nE, np_, files)[2], p1)


def test_ldg_rejects_zero_capacity(O):
    files = [(np.array([0, 1, 2]), np.array([1, 2, 0]))]
    with pytest.raises(N.GreError):
        N.vertex_cut_ldg(3, 3, 2, files)           # 3 / 2 / 2 = 0
    with pytest.raises(ValueError):
        O.ldg_assign(3, 3, 2, files)


def test_c1_agent_counts_np4(O):
    """c1's graph (R-MAT scale 16, seed 1) placed on 4 processes as GRE's np = 4
    PageRank harness would load it, and what its PageRank prints."""
    g, files = c1_files(O, 16, 4)
    s, d, p = N.vertex_cut_ldg(g.nV, g.nE, 4, files)
    st = N.vertex_cut_stats(g.nV, 4, OUT, s, d, p)
    nC, nS, nE = int(st[:, 2].sum()), int(st[:, 1].sum()), int(st[:, 5].sum())
    print(f"c1 np=4: combiners {nC}, scatters {nS}, agents {nC + nS}, edges {nE}, cut {(nC + nS) / nE:.6f}, "
          f"edges per part {st[:, 5].tolist()}")
    assert nE == g.nE
    assert np.array_equal(p, O.ldg_assign(g.nV, g.nE, 4, files)[2])
