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


def test_c4_weighted_ppr_adjoint_f32_vs_f64(N, O):
    d = 0.85
    g = N.Graph.rmat(22, 16, 5, False)
    rng = np.random.default_rng(4)
    nV = g.nV
    p = np.zeros(nV)
    p[rng.choice(nV, 1000, replace=False)] = 1.0 / 1000
    tg = rng.choice(nV, 1024, replace=False).astype(np.uint32)
    res = {}
    for dt in (N.F32, N.F64):
        pr = N.PageRank(g, dt, 1, weighted=True)
        pr.set_edge_weights(None, seed=5)
        pr.set_personalization(p)
        it, r = pr.run(d, 1000, 1e-7 if dt == N.F32 else 1e-13)
        assert r < (1e-7 if dt == N.F32 else 1e-13) and it < 1000
        pr.adjoint(tg, d, 1000, 1e-7 if dt == N.F32 else 1e-13)
        gW, graw, gp = pr.gradients(d)
        res[dt] = (pr.ranks(), graw, gp)
        if dt == N.F64:
            wraw, _ = pr.edge_weights()
    x32, g32, p32 = res[N.F32]
    x64, g64, p64 = res[N.F64]
    assert rel_l1(x32, x64) < 1e-6
    assert rel_l1(g32, g64) < 1e-4
    assert rel_l1(p32, p64) < 1e-4
    # the FP64 forward against the oracle's FP64 power iteration at full size
    og, src = download(N, O, g)
    rs = np.bincount(src, weights=wraw.astype(np.float64), minlength=nV)
    wn = wraw / rs[src]
    xref, _ = O.ppr(og, wn[og.ceid.astype(np.int64)], p, d, tol=1e-14, maxit=2000)
    assert rel_l1(x64, xref) < 1e-11
