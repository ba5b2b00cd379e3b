"""Pins for the oracle (CPU only).

Each check compares the oracle against something that is NOT the oracle's own
formula: hand-derived answers, scipy's graph routines, dense numpy linear
algebra, brute-force path enumeration, or a second algorithm for the same
quantity.  GRE ships no tests or golden vectors for this path, and GRE itself
cannot be built here without MPI and source patches, so these are the pins.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
def tiny_directed(O):
    # 0->1, 0->2, 1->2, 2->0, 3->2 ; vertex 4 isolated
    src = np.array([0, 0, 1, 2, 3], np.uint32)
    dst = np.array([1, 2, 2, 0, 2], np.uint32)
    return O.graph_from_edges(5, src, dst)


def symmetrise(src, dst):
    s = np.concatenate([src, dst])
    d = np.concatenate([dst, src])
    k = np.unique((s.astype(np.uint64) << np.uint64(32)) | d.astype(np.uint64))
    return (k >> np.uint64(32)).astype(np.uint32), (k & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def to_scipy(g, data=None):
    src = np.repeat(np.arange(g.nV), g.outdeg)
    d = np.ones(g.nE) if data is None else data
    return sp.csr_matrix((d, (src, g.idx.astype(np.int64))), shape=(g.nV, g.nV))


def random_graph(O, n, m, seed, sym=False):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, n, m).astype(np.uint32)
    d = rng.integers(0, n, m).astype(np.uint32)
    keep = s != d
    s, d = s[keep], d[keep]
    if sym:
        s, d = symmetrise(s, d)
    else:
        k = np.unique((s.astype(np.uint64) << np.uint64(32)) | d.astype(np.uint64))
        s, d = (k >> np.uint64(32)).astype(np.uint32), (k & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    return O.graph_from_edges(n, s, d)


# ---------------------------------------------------------------------------
# R-MAT generator and CSR/CSC builds
# ---------------------------------------------------------------------------
def test_rmat_quadrant_statistics(O):
    # top-level quadrant of each raw draw follows (a,b,c,d) = (.57,.19,.19,.05)
    S = 12
    u, v = O.rmat_raw(S, seed=7, n=400_000)
    ub, vb = u >> (S - 1), v >> (S - 1)
    frac = [np.mean((ub == 0) & (vb == 0)), np.mean((ub == 0) & (vb == 1)),
            np.mean((ub == 1) & (vb == 0)), np.mean((ub == 1) & (vb == 1))]
    assert np.allclose(frac, [0.57, 0.19, 0.19, 0.05], atol=0.004)
    # every level draws independently: the lowest bit has the same law
    lb_u, lb_v = u & 1, v & 1
    assert abs(np.mean((lb_u == 0) & (lb_v == 0)) - 0.57) < 0.004


def test_rmat_graph_invariants(O):
    g = O.rmat_graph(10, 8, seed=1)
    keys = (np.repeat(np.arange(g.nV, dtype=np.uint64), g.outdeg) << np.uint64(32)) | g.idx
    assert np.all(np.diff(keys.astype(np.int64)) > 0)           # sorted, no duplicates
    src = np.repeat(np.arange(g.nV), g.outdeg)
    assert not np.any(src == g.idx)                             # no self-loops
    u, v = O.rmat_raw(10, seed=1, n=8 << 10)
    raw = set(zip(u.tolist(), v.tolist()))
    raw = {e for e in raw if e[0] != e[1]}
    assert raw == set(zip(src.tolist(), g.idx.tolist()))        # exactly the distinct raw edges


def test_rmat_symmetric(O):
    g = O.rmat_graph(9, 8, seed=2, sym=True)
    a = to_scipy(g)
    assert (a != a.T).nnz == 0
    assert np.array_equal(g.off, g.coff) and np.array_equal(g.idx, g.cidx)


@pytest.mark.parametrize("S", [1, 2, 3, 7, 12, 16])
def test_rmat_scramble_is_a_bijection(O, S):
    x = np.arange(1 << S, dtype=np.uint32)
    y = O.rmat_scramble(S, 6, x)
    assert np.array_equal(np.sort(y), x)
    if S >= 7:
        assert np.mean(y == x) < 0.05                              # it moves almost every label
        assert not np.array_equal(O.rmat_scramble(S, 7, x), y)    # and depends on the seed


@pytest.mark.parametrize("sym", [False, True])
def test_rmat_scrambled_graph_is_the_relabelled_graph(O, sym):
    """The scrambled graph holds exactly the edges (p(u), p(v)) of the plain one."""
    S = 11
    g = O.rmat_graph(S, 8, seed=3, sym=sym)
    h = O.rmat_graph(S, 8, seed=3, sym=sym, scramble=True)
    src = np.repeat(np.arange(g.nV, dtype=np.uint32), g.outdeg)
    pu, pv = O.rmat_scramble(S, 3, src), O.rmat_scramble(S, 3, g.idx)
    want = np.unique((pu.astype(np.uint64) << np.uint64(32)) | pv.astype(np.uint64))
    hsrc = np.repeat(np.arange(h.nV, dtype=np.uint64), h.outdeg)
    got = (hsrc << np.uint64(32)) | h.idx.astype(np.uint64)
    assert np.array_equal(got, want)
    assert int(np.argmax(h.outdeg)) != int(np.argmax(g.outdeg))   # the hub left vertex 0


def test_rmat_fixture(O):
    import os
    f = os.path.join(os.path.dirname(__file__), "golden", "rmat_s10_ef8_seed1.npz")
    z = np.load(f)
    g = O.rmat_graph(10, 8, seed=1)
    assert np.array_equal(g.off, z["off"]) and np.array_equal(g.idx, z["idx"])


def test_csr_csc_input_order(O):
    # GraphInCSR::build/rbuild keep input order inside each row
    src = np.array([2, 0, 2, 1, 0, 2], np.uint32)
    dst = np.array([1, 2, 0, 2, 1, 1], np.uint32)
    g = O.graph_from_edges(3, src, dst)
    assert g.off.tolist() == [0, 2, 3, 6]
    assert g.idx.tolist() == [2, 1, 2, 1, 0, 1]
    # CSC row of 1: sources in CSR-position order: 0 (pos 1), 2 (pos 3), 2 (pos 5)
    assert g.coff.tolist() == [0, 1, 4, 6]
    assert g.cidx.tolist() == [2, 0, 2, 2, 0, 1]
    for p in range(g.nE):
        e = int(g.ceid[p])
        assert g.idx[e] == np.searchsorted(g.coff, p, side="right") - 1


# ---------------------------------------------------------------------------
# GRE engine restatement: hand-derived answers
# ---------------------------------------------------------------------------
def test_gre_pagerank_known_answer(O):
    g = tiny_directed(O)
    st = O.default_state(O.PAGERANK, g.nV)
    ss, counts, _ = O.gre_run(O.PAGERANK, O.DIR_OUT, g, st, max_ss=10)
    pr = st[2]
    assert np.allclose(pr, [0.2775, 0.21375, 0.46875, 0.15, 0.15], rtol=0, atol=1e-15)
    assert ss == 10 and counts.tolist() == [4] * 10


def test_gre_bfs_known_answer(O):
    s, d = symmetrise(np.array([0, 0, 1, 2], np.uint32), np.array([1, 2, 2, 3], np.uint32))
    g = O.graph_from_edges(5, s, d)
    st = O.default_state(O.BFS, g.nV)
    ss, counts, _ = O.gre_run(O.BFS, O.DIR_IN, g, st, srcs=[0])
    assert counts.tolist() == [2, 1, 0] and ss == 2
    m = np.uint64(0xFFFFFFFFFFFFFFFF)
    assert st[2].tolist() == [0, 0, 0, 0, int(m)]


def test_gre_countin_known_answer(O):
    g = tiny_directed(O)
    st = O.default_state(O.COUNTIN, g.nV)
    O.gre_run(O.COUNTIN, O.DIR_OUT, g, st, max_ss=1)
    assert st[2].tolist() == [1, 1, 3, 0, 0]          # in-degrees
    st = O.default_state(O.COUNTIN, g.nV)
    O.gre_run(O.COUNTIN, O.DIR_IN, g, st, max_ss=1)   # the driver's run<In>
    assert st[2].tolist() == [2, 1, 1, 1, 0]          # messages along reverse edges
    assert st[2].sum() == g.outdeg.sum()


# ---------------------------------------------------------------------------
# GRE engine restatement vs independent algorithms
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("seed", [2, 3])
def test_gre_bfs_vs_scipy(O, seed):
    g = O.rmat_graph(11, 16, seed=seed, sym=True)
    src = int(np.argmax(g.outdeg))
    lv = O.bfs_levels(g, src)
    dist = csg.shortest_path(to_scipy(g), unweighted=True, indices=src)
    ref = np.where(np.isinf(dist), -1, dist).astype(np.int32)
    assert np.array_equal(lv, ref)
    st = O.default_state(O.BFS, g.nV)
    ss, counts, _ = O.gre_run(O.BFS, O.DIR_IN, g, st, srcs=[src])
    sizes = np.bincount(lv[lv >= 0])
    assert counts.tolist() == sizes[1:].tolist() + [0]
    reached = lv >= 0
    assert np.all(st[2][reached] == src)
    assert np.all(st[2][~reached] == np.uint64(0xFFFFFFFFFFFFFFFF))


def test_gre_cc_vs_scipy(O):
    g = O.rmat_graph(11, 4, seed=5, sym=True)
    n, lab = csg.connected_components(to_scipy(g), directed=False)
    minv = np.full(n, g.nV, np.int64)
    np.minimum.at(minv, lab, np.arange(g.nV))
    ref = minv[lab]
    uf = O.cc_labels(g)
    assert np.array_equal(uf, ref)
    st = O.default_state(O.CC, g.nV)
    O.gre_run(O.CC, O.DIR_INOUT, g, st)
    assert np.array_equal(st[2], ref.astype(np.uint64))


def test_gre_sssp_vs_scipy(O):
    g = O.rmat_graph(10, 16, seed=2, sym=True)
    w = O.sssp_weights(g, seed=3)
    # one weight per undirected edge, shared by both directions
    src = np.repeat(np.arange(g.nV), g.outdeg)
    wm = {(int(a), int(b)): int(c) for a, b, c in zip(src, g.idx, w)}
    assert all(wm[(b, a)] == c for (a, b), c in wm.items())
    assert w.min() >= 1 and w.max() <= 255
    s0 = int(np.argmax(g.outdeg))
    dref = csg.dijkstra(to_scipy(g, w.astype(np.float64)), indices=s0)
    ref = np.where(np.isinf(dref), 0xFFFFFFFF, dref).astype(np.uint32)
    assert np.array_equal(O.dijkstra(g, w, s0), ref)
    st = O.default_state(O.SSSP, g.nV)
    O.gre_run(O.SSSP, O.DIR_OUT, g, st, aux=w, srcs=[s0])
    assert np.array_equal(st[1], ref)


def test_gre_pagerank_closed_form(O):
    g = O.rmat_graph(10, 8, seed=1)
    st = O.default_state(O.PAGERANK, g.nV)
    ss, counts, _ = O.gre_run(O.PAGERANK, O.DIR_OUT, g, st, max_ss=10)
    contrib = np.where(g.outdeg > 0, 0.15 / np.maximum(g.outdeg, 1), 0.0)
    a = to_scipy(g)
    ref = np.where(g.indeg > 0, 0.15 + 0.85 * (a.T @ contrib), 0.15)
    assert np.allclose(st[2], ref, rtol=1e-13, atol=0)
    assert np.all(counts == np.count_nonzero(g.outdeg))


# ---------------------------------------------------------------------------
# Brandes
# ---------------------------------------------------------------------------
def brute_dependency(g, s):
    """delta_s(v) = sum_t sigma_sv sigma_vt / sigma_st over shortest s-t paths through v."""
    n = g.nV
    a = to_scipy(g).toarray() > 0
    dist = np.full((n, n), -1)
    sig = np.zeros((n, n))
    for x in range(n):
        dist[x, x] = 0
        sig[x, x] = 1
        frontier = [x]
        while frontier:
            nxt = []
            for u in frontier:
                for v in np.nonzero(a[u])[0]:
                    if dist[x, v] < 0:
                        dist[x, v] = dist[x, u] + 1
                        nxt.append(v)
                    if dist[x, v] == dist[x, u] + 1:
                        sig[x, v] += sig[x, u]
            frontier = nxt
    dl = np.zeros(n)
    for v in range(n):
        if v == s or dist[s, v] < 0:
            continue
        for t in range(n):
            if t in (s, v) or dist[s, t] < 0 or dist[v, t] < 0:
                continue
            if dist[s, v] + dist[v, t] == dist[s, t]:
                dl[v] += sig[s, v] * sig[v, t] / sig[s, t]
    return dl, sig[s], dist[s]


@pytest.mark.parametrize("seed", [0, 1])
def test_brandes_vs_brute_force(O, seed):
    g = random_graph(O, 30, 70, seed, sym=(seed == 0))
    for s in (0, 7, 19):
        sig, dl, lv = O.brandes_one(g, s)
        bdl, bsig, bdist = brute_dependency(g, s)
        assert np.array_equal(lv, np.where(bdist < 0, -1, bdist))
        assert np.array_equal(sig, bsig)
        dl_ex = dl.copy()
        dl_ex[s] = 0.0
        assert np.allclose(dl_ex, bdl, rtol=1e-12, atol=1e-12)


def test_gre_bc_pipeline_vs_brandes(O):
    """usr/test_bc.cpp:61-110: Forward run<Out>, then Backward run<In> per level."""
    g = O.rmat_graph(9, 8, seed=4, sym=True)
    src = int(np.argmax(g.outdeg))
    fst = O.default_state(O.BC_FORWARD, g.nV)
    ss, counts, level = O.gre_run(O.BC_FORWARD, O.DIR_OUT, g, fst, srcs=[src], level=1)
    vtx = fst[2]
    sig, dl, lv = O.brandes_one(g, src)
    assert np.array_equal(vtx["level"], lv)
    assert np.array_equal(vtx["numPaths"][lv >= 0], sig[lv >= 0].astype(np.int64).astype(np.int32))
    nlev = level
    bst = O.default_state(O.BC_BACKWARD, g.nV)
    for i in range(nlev - 2, -1, -1):
        O.gre_run(O.BC_BACKWARD, O.DIR_IN, g, bst, aux=vtx, level=i)
    dep = bst[2]
    assert np.allclose(dep[lv >= 0], dl[lv >= 0], rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------------------
# PageRank / PPR / adjoint
# ---------------------------------------------------------------------------
def dense_google(g, d):
    n = g.nV
    A = to_scipy(g).toarray()
    deg = g.outdeg.astype(np.float64)
    P = np.zeros((n, n))
    nz = deg > 0
    P[nz] = A[nz] / deg[nz, None]
    dang = (~nz).astype(np.float64)
    return d * (P.T + np.outer(np.full(n, 1.0 / n), dang)) + (1 - d) / n * np.ones((n, n))


def test_pagerank_vs_dense_google_matrix(O):
    g = random_graph(O, 60, 200, 3)
    G = dense_google(g, 0.85)
    x = np.full(g.nV, 1.0 / g.nV)
    for _ in range(20):
        x = G @ x
    assert np.allclose(O.pagerank(g, 0.85, 20), x, rtol=1e-13, atol=0)
    assert abs(O.pagerank(g, 0.85, 20).sum() - 1.0) < 1e-13


def ppr_setup(O, n=50, m=180, seed=11):
    g = random_graph(O, n, m, seed)
    rng = np.random.default_rng(seed)
    wraw = np.exp(rng.standard_normal(g.nE))
    src = np.repeat(np.arange(g.nV), g.outdeg)
    rs = np.bincount(src, weights=wraw, minlength=g.nV)
    wr = wraw / rs[src]
    wc = wr[g.ceid.astype(np.int64)]
    p = np.zeros(g.nV)
    p[rng.choice(g.nV, 5, replace=False)] = 0.2
    return g, wraw, rs, wr, wc, p, src


def test_ppr_vs_linear_solve(O):
    g, wraw, rs, wr, wc, p, src = ppr_setup(O)
    d = 0.85
    W = np.zeros((g.nV, g.nV))
    W[src, g.idx] = wr
    gd = (g.outdeg == 0).astype(np.float64)
    M = d * (W.T + np.outer(p, gd))
    xref = np.linalg.solve(np.eye(g.nV) - M, (1 - d) * p)
    x, it = O.ppr(g, wc, p, d, tol=1e-15, maxit=5000)
    assert np.allclose(x, xref, rtol=1e-11, atol=1e-14)


def test_ppr_adjoint_and_gradients(O):
    g, wraw, rs, wr, wc, p, src = ppr_setup(O)
    d = 0.85
    tg = np.array([1, 4, 9, 20, 33])
    c = np.zeros(g.nV)
    c[tg] = 1.0
    W = np.zeros((g.nV, g.nV))
    W[src, g.idx] = wr
    gd = (g.outdeg == 0).astype(np.float64)
    M = d * (W.T + np.outer(p, gd))
    lref = np.linalg.solve((np.eye(g.nV) - M).T, c)
    lam, _ = O.ppr_adjoint(g, wr, p, c, d, tol=1e-15, maxit=5000)
    assert np.allclose(lam, lref, rtol=1e-11, atol=1e-13)
    x, _ = O.ppr(g, wc, p, d, tol=1e-15, maxit=5000)
    gW, graw = O.ppr_edge_grad(g, wr, rs, x, lam, d)

    def loss(wr_new, p_new):
        wc_new = wr_new[g.ceid.astype(np.int64)]
        xx, _ = O.ppr(g, wc_new, p_new, d, tol=1e-15, maxit=5000)
        return xx[tg].sum()

    rng = np.random.default_rng(0)
    for e in rng.choice(g.nE, 12, replace=False):
        # d/dW_e (normalised weight, others fixed)
        h = 1e-6
        wp, wm = wr.copy(), wr.copy()
        wp[e] += h
        wm[e] -= h
        fd = (loss(wp, p) - loss(wm, p)) / (2 * h)
        assert abs(fd - gW[e]) < 1e-7 * max(1.0, abs(fd))
        # d/dw_raw through the row normalisation
        u = src[e]
        rows = slice(int(g.off[u]), int(g.off[u + 1]))

        def wr_from(wraw_new):
            out = wr.copy()
            out[rows] = wraw_new[rows] / wraw_new[rows].sum()
            return out

        hp, hm = wraw.copy(), wraw.copy()
        hr = 1e-6 * wraw[e]
        hp[e] += hr
        hm[e] -= hr
        fd2 = (loss(wr_from(hp), p) - loss(wr_from(hm), p)) / (2 * hr)
        assert abs(fd2 - graw[e]) < 1e-6 * max(1.0, abs(fd2))
    # d/dp: ((1-d) + d g.x) lambda (p moved freely, no renormalisation)
    j = int(np.nonzero(p)[0][0])
    hp = p.copy()
    hp[j] += 1e-6
    hm = p.copy()
    hm[j] -= 1e-6
    fdp = (loss(wr, hp) - loss(wr, hm)) / 2e-6
    gpj = ((1 - d) + d * x[g.outdeg == 0].sum()) * lam[j]
    assert abs(fdp - gpj) < 1e-7 * max(1.0, abs(fdp))


# ---------------------------------------------------------------------------
# PageRank_delta (include/Kernels/PageRank.hpp:129-234)
# ---------------------------------------------------------------------------
def delta_graph(O, n=110, sinks=10, m=650, seed=21):
    """A random digraph on n vertices plus `sinks` vertices that have in-edges
    but no out-edges (so the out-degree guard of the activation matters)."""
    rng = np.random.default_rng(seed)
    s = rng.integers(0, n, m)
    d = rng.integers(0, n, m)
    keep = s != d
    s, d = s[keep], d[keep]
    s = np.concatenate([s, rng.integers(0, n, 3 * sinks)])
    d = np.concatenate([d, n + np.arange(3 * sinks) % sinks])
    k = np.unique((s.astype(np.uint64) << np.uint64(32)) | d.astype(np.uint64))
    s, d = (k >> np.uint64(32)).astype(np.uint32), (k & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    return O.graph_from_edges(n + sinks, s, d)


def transition(g):
    """0/1 adjacency B and row-stochastic P (rows of out-degree 0 are zero)."""
    A = to_scipy(g).toarray()
    deg = g.outdeg.astype(np.float64)
    P = np.zeros_like(A)
    nz = deg > 0
    P[nz] = A[nz] / deg[nz, None]
    return (A > 0).astype(np.int64), P


def test_gre_pagerank_delta_matrix_powers(O):
    """Away from the 1e-5 threshold, superstep k adds delta_k = (0.85 P^T)^(k+1) 0.15 1
    to pr, and the vertices active after it are those with out-edges at the end of
    some walk of length k+1 (numpy matrix powers; no engine rule is restated)."""
    g = delta_graph(O)
    B, P = transition(g)
    assert np.count_nonzero(g.outdeg == 0) >= 10
    y = np.full(g.nV, 0.15)
    pr = y.copy()
    walk = np.ones(g.nV, np.int64)
    for k in range(8):
        y = 0.85 * (P.T @ y)
        walk = (B.T @ walk > 0).astype(np.int64)
        assert np.array_equal(y > 0, walk > 0)
        assert y[y > 0].min() > 2e-5            # the regime this closed form holds in
        pr += y
        st = O.default_state(O.PAGERANK_DELTA, g.nV)
        ss, counts, _ = O.gre_run(O.PAGERANK_DELTA, O.DIR_OUT, g, st, max_ss=k + 1)
        assert ss == k + 1
        assert np.allclose(st[2], pr, rtol=1e-13, atol=0)
        assert counts[k] == np.count_nonzero((walk > 0) & (g.outdeg > 0))


def test_gre_pagerank_delta_run_to_halt(O):
    """The full run halts with every vertex's last delta at or below 1e-5. The
    trajectory (pr after each superstep) splits into the deltas that propagated and
    those dropped; the fixpoint x* = 0.15 (I - 0.85 P^T)^-1 1 then satisfies
    x* - pr = (I - 0.85 P^T)^-1 0.85 P^T D exactly, D = the dropped deltas, and
    each superstep's count is the number of vertices with out-edges and delta > 1e-5."""
    g = delta_graph(O, seed=5)
    _, P = transition(g)
    I = np.eye(g.nV)
    st = O.default_state(O.PAGERANK_DELTA, g.nV)
    ss, counts, _ = O.gre_run(O.PAGERANK_DELTA, O.DIR_OUT, g, st)
    assert 5 < ss < 1024 and counts[-1] == 0
    prev = np.full(g.nV, 0.15)
    D = np.zeros(g.nV)
    for k in range(ss + 1):
        sk = O.default_state(O.PAGERANK_DELTA, g.nV)
        O.gre_run(O.PAGERANK_DELTA, O.DIR_OUT, g, sk, max_ss=k + 1)
        delta = sk[2] - prev
        prev = sk[2]
        assert delta.min() >= 0.0
        live = (delta > 1e-5 * (1 + 1e-9)) & (g.outdeg > 0)
        dead = (delta > 0) & ~live
        assert not np.any((np.abs(delta - 1e-5) < 1e-14) & (g.outdeg > 0))   # no knife edge
        assert counts[k] == np.count_nonzero(live)
        D += np.where(dead, delta, 0.0)
    assert np.array_equal(prev, st[2])
    xstar = np.linalg.solve(I - 0.85 * P.T, np.full(g.nV, 0.15))
    gap = np.linalg.solve(I - 0.85 * P.T, 0.85 * (P.T @ D))
    assert np.allclose(xstar - st[2], gap, rtol=1e-9, atol=1e-13)
    assert np.all(xstar - st[2] > -1e-13)


@pytest.mark.parametrize("deg,counts", [(76500, [0]), (76501, [0]), (76499, [76499, 0])])
def test_gre_pagerank_delta_threshold(O, deg, counts):
    """Six hubs point at the same `deg` targets, every target points at one sink.
    A target's combine adds six messages 0.15/deg; at deg = 76500 that sum times
    0.85 is exactly 1e-5 in IEEE double, and GRE's strict `delta > err`
    (PageRank.hpp:182) leaves the targets inactive; one fewer target lifts the
    delta above 1e-5. The sink gets a large delta but has no out-edges."""
    m = 0.15 / deg
    acc = 0.0
    for _ in range(6):
        acc += m
    assert (0.85 * acc == 1e-5) == (deg == 76500)
    assert (0.85 * acc > 1e-5) == (deg == 76499)
    hubs = np.repeat(np.arange(6), deg)
    tg = np.tile(np.arange(6, 6 + deg), 6)
    sink = 6 + deg
    src = np.concatenate([hubs, np.arange(6, 6 + deg)]).astype(np.uint32)
    dst = np.concatenate([tg, np.full(deg, sink)]).astype(np.uint32)
    g = O.graph_from_edges(sink + 1, src, dst)
    st = O.default_state(O.PAGERANK_DELTA, g.nV)
    ss, c, _ = O.gre_run(O.PAGERANK_DELTA, O.DIR_OUT, g, st)
    assert c.tolist() == counts and ss == len(counts) - 1
    assert st[2][6] == 0.15 + 0.85 * acc


# ---------------------------------------------------------------------------
# CPU baseline: the multi-threaded GRE-style superstep
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("threads", [1, 4])
def test_gre_pagerank_mt_equals_pagerank(O, threads):
    """The timed CPU baseline computes the same FP64 power iteration as o_pagerank
    (itself pinned by the dense Google matrix above), up to summation order."""
    g = O.rmat_graph(12, 16, seed=1)
    x, secs = O.gre_pagerank_mt(g, nthreads=threads, iters=20)
    ref = O.pagerank(g, 0.85, 20)
    assert secs > 0
    assert np.abs(x - ref).sum() / ref.sum() < 1e-13
    assert abs(x.sum() - 1.0) < 1e-12


def _file_order(g, np_):
    """Edge files of the np-process ingress: file i = CSR rows u with u % np == i."""
    src = np.repeat(np.arange(g.nV, dtype=np.uint32), g.outdeg)
    files = []
    for i in range(np_):
        m = (src % np_) == i
        files.append((src[m], g.idx[m]))
    return files


@pytest.mark.parametrize("np_,workers", [(1, 3), (2, 2), (4, 3), (3, 1)])
def test_gre_pagerank_np_equals_pagerank(O, np_, workers):
    """The np-process CPU baseline: the same FP64 ranks as o_pagerank up to
    summation order; each superstep sends one scatter message per scatter agent
    and one combiner message per combiner agent of the LDG cut, whose counts the
    vertex-cut statistics give independently; every edge lands in one part."""
    g = O.rmat_graph(11, 16, seed=1)
    x, secs, info = O.gre_pagerank_np(g, np_=np_, workers=workers, iters=20)
    ref = O.pagerank(g, 0.85, 20)
    assert secs > 0
    assert np.abs(x - ref).sum() / ref.sum() < 1e-13
    src, dst, part = O.ldg_assign(g.nV, g.nE, np_, _file_order(g, np_))
    st = O.vertex_cut_stats(g.nV, np_, O.DIR_OUT, src, dst, part)
    assert info["scatter_msgs"] == st[:, 1].sum()      # scatters (Out direction)
    assert info["combiner_msgs"] == st[:, 2].sum()     # combiners
    assert np.array_equal(info["part_edges"], st[:, 5])
    assert info["part_edges"].sum() == g.nE
    if np_ == 1:
        assert info["scatter_msgs"] == 0 and info["combiner_msgs"] == 0



def test_gre_pagerank_np_worker_list(O):
    """One ingress, one run per worker count: every run starts afresh and ends
    at the same ranks."""
    g = O.rmat_graph(10, 16, seed=2)
    ref = O.pagerank(g, 0.85, 10)
    for ws in ([1, 3], [3, 1]):
        x, secs, info = O.gre_pagerank_np(g, np_=4, workers=ws, iters=10)
        assert len(secs) == 2 and min(secs) > 0
        assert np.abs(x - ref).sum() / ref.sum() < 1e-13
