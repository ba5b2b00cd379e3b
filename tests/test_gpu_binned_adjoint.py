"""GPU parity of the binned single-vector adjoint (FP32 handles on the binned
schedule) through the C ABI, against the FP64 oracle (oracle/gre_oracle.c
o_ppr_adjoint: lambda_0 = c, then lambda'_u = c_u + d (sum_{u->v} w_uv
lambda_v + [outdeg u == 0] p.lambda)).

The adjoint runs the forward's binned layout over the reversed graph: the
entries of one chunk to one row are summed in FP32 (at most 16), each row's
partial sums are added as 64-bit fixed point at a scale set from max |lambda|
of the iteration before, so results are deterministic.  Tolerances: 1e-6
relative L1 against the FP64 oracle for cotangents of one sign (the bar of the
FP32 forward), 2e-6 for signed cotangents (whose sums cancel), and 1e-6 against
the tile schedule's FP32 adjoint.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BIN, TILE = 8, 1
D = 0.85


def rel_l1(a, b):
    b = np.asarray(b, np.float64)
    return float(np.abs(np.asarray(a, np.float64) - b).sum() / max(np.abs(b).sum(), 1e-300))


def norm_weights(og, wraw):
    """The handle's normalised CSR weights: w_uv / sum_v w_uv, or 1 / outdeg(u) unweighted."""
    src = np.repeat(np.arange(og.nV), og.outdeg)
    if wraw is None:
        return 1.0 / og.outdeg[src].astype(np.float64)
    wr = wraw.astype(np.float64)
    rs = np.bincount(src, weights=wr, minlength=og.nV)
    return np.where(rs[src] > 0, wr / rs[src], 0.0)


def handle(N, g, sched, weighted, p):
    h = N.PageRank(g, N.F32, 1, weighted=weighted)
    h.set_schedule(sched, 8)
    if weighted:
        h.set_edge_weights(None, seed=5)
    if p is not None:
        h.set_personalization(p)
    return h


@pytest.mark.parametrize("scale,weighted,with_p,signed", [
    (12, False, False, True), (12, True, True, False), (18, False, True, True), (18, True, False, True),
    (18, True, True, False)])
def test_binned_adjoint_matches_oracle(N, O, scale, weighted, with_p, signed):
    """Several bins at scale 18 (the vertices with out-edges are the rows), rows
    without out-edges (the dangling term), uniform or sparse personalisation,
    weighted or 1 / out-degree."""
    g = N.Graph.rmat(scale, 16, 3, False)
    og = O.rmat_graph(scale, 16, 3, False)
    rng = np.random.default_rng(scale)
    p = None
    if with_p:
        p = np.zeros(og.nV)
        p[rng.choice(og.nV, 40, replace=False)] = 1.0 / 40
    c = rng.standard_normal(og.nV) if signed else rng.random(og.nV)
    bh = handle(N, g, BIN, weighted, p)
    th = handle(N, g, TILE, weighted, p)
    for h in (bh, th):
        h.run(D, 10)
        it, _ = h.adjoint_cotangent(c.astype(np.float32), D, 40)
        assert it == 40
    assert bh.schedule()[0] == "bin"
    wraw = bh.edge_weights()[0] if weighted else None
    pref = p if p is not None else np.full(og.nV, 1.0 / og.nV)
    lam_ref, _ = O.ppr_adjoint(og, norm_weights(og, wraw), pref, c.astype(np.float32).astype(np.float64), D,
                               tol=0.0, maxit=40)
    lam = bh.adjoint_vector()
    assert rel_l1(lam, lam_ref) < (2e-6 if signed else 1e-6)
    assert rel_l1(lam, th.adjoint_vector()) < 1e-6
    # the gradients read lambda and (W lambda) back in vertex order
    gW, graw, gp = bh.gradients(D)
    rW, rraw, rp = th.gradients(D)
    assert rel_l1(gW, rW) < 1e-5
    assert rel_l1(gp, rp) < 1e-5
    if weighted:
        assert rel_l1(graw, rraw) < 1e-5


def test_binned_adjoint_is_deterministic(N):
    """Two handles, and two solves on one handle, give the same bits; the
    residual is the last iteration's L1 change."""
    g = N.Graph.rmat(16, 16, 8, False)
    c = np.random.default_rng(1).standard_normal(g.nV).astype(np.float32)
    a = handle(N, g, BIN, True, None)
    b = handle(N, g, BIN, True, None)
    for h in (a, b):
        h.run(D, 5)
    a.adjoint_cotangent(c, D, 29)
    l29 = a.adjoint_vector().astype(np.float64)
    _, res = a.adjoint_cotangent(c, D, 30)
    l30 = a.adjoint_vector()
    b.adjoint_cotangent(c, D, 30)
    assert np.array_equal(l30, b.adjoint_vector())
    a.adjoint_cotangent(c, D, 30)
    assert np.array_equal(l30, a.adjoint_vector())
    ref = np.abs(l30 - l29).sum()
    assert abs(res - ref) <= 1e-5 * ref


def test_binned_adjoint_converged_run(N, O):
    """A run to a tolerance stops early and lands on the converged oracle."""
    g = N.Graph.rmat(15, 16, 4, False)
    og = O.rmat_graph(15, 16, 4, False)
    tg = np.random.default_rng(4).choice(og.nV, 16, replace=False).astype(np.uint32)
    h = handle(N, g, BIN, False, None)
    h.run(D, 5)
    it, res = h.adjoint(tg, D, 1000, 1e-7)
    assert 10 < it < 1000 and res < 1e-7
    c = np.zeros(og.nV)
    c[tg] = 1.0
    lam_ref, _ = O.ppr_adjoint(og, norm_weights(og, None), np.full(og.nV, 1.0 / og.nV), c, D, tol=1e-13,
                               maxit=5000)
    assert rel_l1(h.adjoint_vector(), lam_ref) < 1e-5


def test_binned_adjoint_multi_slice_bin(N, O):
    """16,384 vertices with 2,000,000 out-edges in all, vertex u -> the
    out-degree-0 vertex 16,384 + s for every s = u mod 16,384: in the reversed
    graph the 16,384 vertices form one bin, fed by 2,000,000 sources that each
    reach it once per chunk, so no entries share a slot and the bin's slots
    take two accumulate slices, whose sums meet in 64-bit integers in HBM."""
    nd, ns = 16384, 2_000_000
    n = nd + ns
    src = (np.arange(ns, dtype=np.uint64) % nd).astype(np.uint32)
    dst = np.arange(nd, n, dtype=np.uint32)
    g = N.Graph.from_edges(n, src, dst)
    og = O.graph_from_edges(n, src, dst)
    c = np.random.default_rng(9).standard_normal(n).astype(np.float32)
    h = handle(N, g, BIN, False, None)
    h.run(D, 5)
    it, _ = h.adjoint_cotangent(c, D, 20)
    assert it == 20
    lam_ref, _ = O.ppr_adjoint(og, norm_weights(og, None), np.full(n, 1.0 / n), c.astype(np.float64), D,
                               tol=0.0, maxit=20)
    lam = h.adjoint_vector()
    assert rel_l1(lam, lam_ref) < 2e-6
    for v in (0, 1, 5000, nd - 1, nd, n - 1):
        assert abs(float(lam[v]) - lam_ref[v]) <= 1e-5 * abs(lam_ref[v]) + 1e-6, v


@pytest.mark.parametrize("weighted", [False, True])
def test_binned_adjoint_scales_exactly(N, weighted):
    """The adjoint is linear in c, and every step (FP32 group sums, the
    fixed-point scale taken from max |lambda|, the integer sums, the division,
    the update, the FP64 p.lambda) commutes with a power-of-two factor while
    no value leaves the normal FP32 range: lambda(2^k c) = 2^k lambda(c) bit
    for bit, for k = 60 and -60.  A fixed-point scale that did not follow
    |lambda| would overflow at 2^60 or lose every bit at 2^-60."""
    g = N.Graph.rmat(14, 16, 2, False)
    c = np.random.default_rng(7).standard_normal(g.nV).astype(np.float32)
    h = handle(N, g, BIN, weighted, None)
    h.run(D, 5)
    h.adjoint_cotangent(c, D, 20)
    lam = h.adjoint_vector()
    assert np.abs(lam).max() > 0
    for k in (60, -60):
        f = np.float32(2.0 ** k)
        h.adjoint_cotangent(c * f, D, 20)
        assert np.array_equal(h.adjoint_vector(), lam * f), k
