"""GPU parity at the full sizes of the measured configurations.

  scale 22 symmetrised, seed 2, SSSP weights U[1,255] seed 3:
      CC labels (GRE's CC program), BFS levels and SSSP distances from sources
      in the largest component, 64-source bitmask BFS;
  scale 20 symmetrised, seed 4: 256-source FP64 Brandes BC, 64-source BFS;
  scale 22 directed, seed 5: weighted personalised PageRank (FP32 forward to
      1e-7 L1, FP64 residual), its adjoint and gradients against FP64.

The device graph is downloaded and the oracle runs on it (the device R-MAT
itself is pinned against the oracle's R-MAT at smaller scales); structural
invariants of the download are checked first.  Integers are bit-exact; BC
dependencies agree to 1e-9 relative, path counts exactly; FP32 PPR ranks to
1e-6 and FP32 gradients to 1e-4 relative L1 of FP64.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_l1(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).sum() / max(np.abs(b).sum(), 1e-300))


def download(N, O, g):
    d = g.download(csc=False)
    og = O.Graph(g.nV, d["off"], d["idx"])
    # rows sorted, no self-loops, no duplicates
    src = np.repeat(np.arange(og.nV, dtype=np.int64), og.outdeg)
    idx = og.idx.astype(np.int64)
    key = src * og.nV + idx
    assert np.all(np.diff(key) > 0)
    assert not np.any(src == idx)
    return og, src


@pytest.fixture(scope="module")
def c2(N, O):
    g = N.Graph.rmat(22, 16, 2, True)
    og, src = download(N, O, g)
    # symmetric: the reverse of every edge is stored
    rkey = np.sort(og.idx.astype(np.int64) * og.nV + src)
    assert np.array_equal(rkey, src * og.nV + og.idx.astype(np.int64))
    lab = O.cc_labels(og)
    big = np.bincount(lab).argmax()
    rng = np.random.default_rng(22)
    srcs = rng.choice(np.nonzero(lab == big)[0], 64, replace=False).astype(np.uint32)
    return g, og, lab, srcs


def test_c2_cc_labels(N, O, c2):
    g, og, lab, _ = c2
    e = N.Engine(g, N.CC)
    e.run(N.DIR_INOUT)
    assert np.array_equal(e.download(N.ARR_VERTEX).astype(np.uint32), lab)


def test_c2_bfs_levels(N, O, c2):
    g, og, _, srcs = c2
    for s in srcs[:4]:
        lev, edges, depth = N.bfs(g, int(s))
        ref = O.bfs_levels(og, int(s))
        assert np.array_equal(lev, ref)
        assert edges == int(og.outdeg[ref >= 0].sum())
    # the 64-source bitmask BFS equals the plain BFS for every source
    mlev, medges = N.msbfs(g, srcs)
    total = 0
    for i, s in enumerate(srcs):
        lev, edges, _ = N.bfs(g, int(s))
        assert np.array_equal(mlev[i], lev)
        total += edges
    assert medges == total


def test_c2_sssp_distances(N, O, c2):
    g, og, _, srcs = c2
    g.set_weights_u32(None, seed=3)
    w = O.sssp_weights(og, seed=3)
    assert np.array_equal(g.weights_u32(), w)
    s = int(srcs[0])
    e = N.Engine(g, N.SSSP)
    e.run(N.DIR_OUT, srcs=[s])
    assert np.array_equal(e.download(N.ARR_VERTEX), O.dijkstra(og, w, s))


def test_c3_bc_256_sources(N, O):
    g = N.Graph.rmat(20, 16, 4, True)
    og, _ = download(N, O, g)
    rng = np.random.default_rng(20)
    srcs = rng.choice(og.nV, 256, replace=False).astype(np.uint32)
    bc, _, _, _ = N.bc(g, srcs)
    # the sum over 256 sources is the sum of four 64-source batches
    parts = sum(N.bc(g, srcs[64 * i:64 * (i + 1)])[0] for i in range(4))
    assert np.abs(bc - parts).sum() <= 1e-12 * np.abs(parts).sum()
    # per-source path counts, dependencies and levels against FP64 Brandes
    pick = srcs[[0, 63, 64, 255]]
    _, sig, dl, lv = N.bc(g, pick, want_parts=True)
    for i, s in enumerate(pick):
        osig, odl, olv = O.brandes_one(og, int(s))
        assert np.array_equal(lv[i], olv)
        assert np.array_equal(sig[i], osig)
        assert rel_l1(dl[i], odl) < 1e-9
    # 64-source BFS at this size against the oracle
    mlev, _ = N.msbfs(g, srcs[:64])
    for i in (0, 31, 63):
        assert np.array_equal(mlev[i], O.bfs_levels(og, int(srcs[i])))


@pytest.fixture(scope="module")
def c4(N, O):
    """Scale 22 directed, seed 5: lognormal(0,1) weights (seed 5) row-normalised,
    p uniform on 1,000 seeds, loss = sum of ranks over 1,024 targets.  Device
    FP32 and FP64 forward to 1e-7 / 1e-13 L1, adjoint and gradients, and the
    oracle's FP64 forward, adjoint and edge gradients on the downloaded graph."""
    d = 0.85
    g = N.Graph.rmat(22, 16, 5, False)
    rng = np.random.default_rng(4)
    nV = g.nV
    p = np.zeros(nV)
    p[rng.choice(nV, 1000, replace=False)] = 1.0 / 1000
    tg = rng.choice(nV, 1024, replace=False).astype(np.uint32)
    res = {}
    for dt in (N.F32, N.F64):
        tol = 1e-7 if dt == N.F32 else 1e-13
        pr = N.PageRank(g, dt, 1, weighted=True)
        pr.set_edge_weights(None, seed=5)
        pr.set_personalization(p)
        it, r = pr.run(d, 1000, tol)
        assert r < tol and it < 1000
        pr.adjoint(tg, d, 1000, tol)
        gW, graw, gp = pr.gradients(d)
        res[dt] = dict(x=pr.ranks().astype(np.float64), lam=pr.adjoint_vector().astype(np.float64), gW=gW,
                       graw=graw, gp=gp.astype(np.float64))
        if dt == N.F64:
            wraw, _ = pr.edge_weights()
        pr.close()
    og, src = download(N, O, g)
    wraw = wraw.astype(np.float64)
    rs = np.bincount(src, weights=wraw, minlength=nV)
    wn = wraw / rs[src]
    xref, _ = O.ppr(og, wn[og.ceid.astype(np.int64)], p, d, tol=1e-14, maxit=2000)
    c = np.zeros(nV)
    c[tg] = 1.0
    lref, _ = O.ppr_adjoint(og, wn, p, c, d, tol=1e-10, maxit=2000)
    gWref, grawref = O.ppr_edge_grad(og, wn, rs, xref, lref, d)
    gpref = ((1 - d) + d * xref[og.outdeg == 0].sum()) * lref
    ref = dict(x=xref, lam=lref, gW=gWref, graw=grawref, gp=gpref)
    return dict(g=g, og=og, p=p, tg=tg, d=d, wraw=wraw, res=res, ref=ref)


def test_c4_f32_vs_f64(N, c4):
    r32, r64 = c4["res"][N.F32], c4["res"][N.F64]
    assert rel_l1(r32["x"], r64["x"]) < 1e-6
    assert rel_l1(r32["graw"], r64["graw"]) < 1e-4
    assert rel_l1(r32["gp"], r64["gp"]) < 1e-4


def test_c4_f64_vs_oracle(N, c4):
    # forward and adjoint against the oracle's FP64 power iterations (pinned by
    # linear solves), gradients against its edge gradients (pinned by central
    # differences); the device returns gradients in FP32
    r64, ref = c4["res"][N.F64], c4["ref"]
    assert rel_l1(r64["x"], ref["x"]) < 1e-11
    assert rel_l1(r64["lam"], ref["lam"]) < 1e-11
    assert rel_l1(r64["gp"], ref["gp"]) < 1e-11
    assert rel_l1(r64["gW"], ref["gW"]) < 1e-5
    assert rel_l1(r64["graw"], ref["graw"]) < 1e-5
    r32 = c4["res"][N.F32]
    assert rel_l1(r32["lam"], ref["lam"]) < 1e-5
    assert rel_l1(r32["graw"], ref["graw"]) < 1e-4


def test_c4_central_differences_64_edges(N, c4):
    """d(loss)/d(raw weight) of 64 sampled edges against central differences of
    the FP64 forward, h = 1e-3 of the weight: 32 edges drawn from the 1 % with the
    largest gradients (where the loss moves measurably), 32 uniformly."""
    g, og, p, tg, d = c4["g"], c4["og"], c4["p"], c4["tg"], c4["d"]
    wraw = c4["wraw"]
    graw = c4["res"][N.F64]["graw"].astype(np.float64)
    rng = np.random.default_rng(64)
    top = np.argsort(-np.abs(graw))[: max(og.nE // 100, 32)]
    edges = np.concatenate([rng.choice(top, 32, replace=False), rng.choice(og.nE, 32, replace=False)])
    pr = N.PageRank(g, N.F64, 1, weighted=True)
    pr.set_personalization(p)
    w32 = wraw.astype(np.float32)
    worst = 0.0
    for e in edges:
        h = 1e-3 * float(wraw[e])
        w_hi, w_lo = np.float32(wraw[e] + h), np.float32(wraw[e] - h)
        vals = []
        for wv in (w_hi, w_lo):
            w2 = w32.copy()
            w2[e] = wv
            pr.set_edge_weights(w2)
            pr.run(d, 2000, 1e-15)
            vals.append(pr.ranks()[tg].sum())
        # the perturbed weights are float32: the step is their difference
        fd = (vals[0] - vals[1]) / (float(w_hi) - float(w_lo))
        err = abs(fd - graw[e])
        worst = max(worst, err / max(abs(fd), 1e-12))
        assert err <= 1e-4 * abs(fd) + 1e-11, (int(e), fd, graw[e])
    pr.close()
    print(f"c4 central differences: worst relative error {worst:.2e}")


def test_c5_batched_ppr_adjoint_32_vectors(N, O):
    """Scale 26 directed (seed 6): 32 personalised vectors, each uniform on 100
    random vertices, in one batched run (V x k = 2^31 values per array), and the
    adjoint of the sum over 1,024 targets on the same batch.  Checks: every
    column sums to 1; the forward's FP64 residual is at the FP32 floor; the
    duality 1_T' x_j = (1 - d) p_j' lambda_j holds per column (a wrong adjoint,
    e.g. one missing the dangling term, fails it); three columns equal the
    single-vector (partitioned) forward and adjoint."""
    d, k, its = 0.85, 32, 150
    g = N.Graph.rmat(26, 16, 6, False)
    nV = g.nV
    rng = np.random.default_rng(6)
    seeds = [rng.choice(nV, 100, replace=False) for _ in range(k)]
    tg = rng.choice(nV, 1024, replace=False).astype(np.uint32)
    P = np.zeros((nV, k), np.float32)
    for j in range(k):
        P[seeds[j], j] = 1.0 / 100
    bp = N.PageRank(g, N.F32, k, weighted=True)
    bp.set_edge_weights(None, seed=6)
    bp.set_personalization(P)
    del P
    it, res = bp.run(d, its)
    assert it == its and res < 1e-6
    X = bp.ranks()
    bp.adjoint(tg, d, its)
    LAM = bp.adjoint_vector()
    bp.close()
    assert X.shape == (nV, k) and LAM.shape == (nV, k)
    colsum = X.sum(axis=0, dtype=np.float64)
    assert np.all(np.abs(colsum - 1.0) < 1e-5), colsum
    lhs = X[tg].astype(np.float64).sum(axis=0)
    rhs = np.array([(1 - d) * LAM[seeds[j], j].astype(np.float64).sum() / 100 for j in range(k)])
    assert np.all(np.abs(lhs - rhs) <= 1e-4 * np.abs(lhs)), (lhs, rhs)
    sp = N.PageRank(g, N.F32, 1, weighted=True)
    sp.set_edge_weights(None, seed=6)
    for j in (0, 17, 31):
        pj = np.zeros(nV, np.float32)
        pj[seeds[j]] = 1.0 / 100
        sp.set_personalization(pj)
        sp.run(d, its)
        assert rel_l1(X[:, j], sp.ranks().astype(np.float64)) < 1e-6
        sp.adjoint(tg, d, its)
        assert rel_l1(LAM[:, j], sp.adjoint_vector().astype(np.float64)) < 1e-6
    assert sp.schedule()[0].startswith(("bin", "part"))   # the autotuned single-vector schedules at this size
    sp.close()
    g.close()
