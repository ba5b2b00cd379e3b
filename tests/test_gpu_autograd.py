"""Personalised PageRank as a differentiable torch op, in the test harness only.

`PPRFunction` wraps the C-ABI forward (gre_pr_run), the cotangent-seeded
adjoint (gre_pr_adjoint_cotangent) and the gradient kernels
(gre_pr_gradients) in a torch.autograd.Function.  Its gradients with respect to
the raw FP32 edge weights and the personalisation vector are checked against
the oracle's FP64 reference gradient and against central finite differences
(h = 1e-3 relative on the weight, 1e-3 absolute on p).

Tolerances: FP64 device gradient vs FP64 oracle gradient 1e-9 relative L1
(the device gradient is stored in FP32 for edges: 1e-6); central differences
agree to 2e-4 relative (the FD truncation error at h = 1e-3).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

D = 0.85
TOL = 1e-15


class PPRFunction(torch.autograd.Function):
    """x = PPR(w_raw, p) on a fixed graph; w_raw in CSR order (FP32), p of shape (nV,) or (nV, k)."""

    @staticmethod
    def forward(ctx, w_raw, p, pr):
        pr.set_edge_weights(w_raw.detach().cpu().numpy().astype(np.float32))
        pr.set_personalization(p.detach().cpu().numpy())
        pr.run(D, 5000, TOL)
        ctx.pr = pr
        return torch.from_numpy(pr.ranks())

    @staticmethod
    def backward(ctx, gx):
        pr = ctx.pr
        pr.adjoint_cotangent(gx.detach().cpu().numpy(), D, 5000, TOL)
        _, graw, gp = pr.gradients(D)
        return torch.from_numpy(graw.astype(np.float64)), torch.from_numpy(gp), None


def rel_l1(a, b):
    a = np.asarray(a, np.float64)
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def problem(N, O, scale=11, seed=5, nseeds=20, k=1):
    g = N.Graph.rmat(scale, 16, seed, False)
    og = O.rmat_graph(scale, 16, seed, False)
    rng = np.random.default_rng(seed)
    shape = (og.nV,) if k == 1 else (og.nV, k)
    p = np.zeros(shape)
    for j in range(k):
        col = rng.choice(og.nV, nseeds, replace=False)
        if k == 1:
            p[col] = 1.0 / nseeds
        else:
            p[col, j] = 1.0 / nseeds
    init = N.PageRank(g, N.F64, k, weighted=True)
    init.set_edge_weights(None, seed=seed)
    wraw, _ = init.edge_weights()
    return g, og, wraw, p


def oracle_grads(O, og, wraw, p, cot):
    """FP64 reference: forward, adjoint from the cotangent, edge and p gradients."""
    src = np.repeat(np.arange(og.nV), og.outdeg)
    wr = wraw.astype(np.float64)
    rs = np.bincount(src, weights=wr, minlength=og.nV)
    wn = np.where(rs[src] > 0, wr / rs[src], 0.0)
    x, _ = O.ppr(og, wn[og.ceid.astype(np.int64)], p, D, tol=TOL, maxit=5000)
    lam, _ = O.ppr_adjoint(og, wn, p, cot(x), D, tol=TOL, maxit=5000)
    _, graw = O.ppr_edge_grad(og, wn, rs, x, lam, D)
    gp = ((1 - D) + D * x[og.outdeg == 0].sum()) * lam
    return x, graw, gp


def test_autograd_sum_of_targets_matches_targets_entry_point(N, O):
    g, og, wraw, p = problem(N, O)
    tg = np.random.default_rng(2).choice(og.nV, 32, replace=False)
    # one fixed tile size: the FP64 summation order then matches between handles
    pr = N.PageRank(g, N.F64, 1, weighted=True)
    pr.set_schedule(1, 8)
    w = torch.tensor(wraw.astype(np.float64), requires_grad=True)
    pt = torch.tensor(p, requires_grad=True)
    x = PPRFunction.apply(w, pt, pr)
    x[torch.from_numpy(tg)].sum().backward()
    # the targets entry point seeds the same cotangent (ones on the targets)
    ref = N.PageRank(g, N.F64, 1, weighted=True)
    ref.set_schedule(1, 8)
    ref.set_edge_weights(wraw)
    ref.set_personalization(p)
    ref.run(D, 5000, TOL)
    ref.adjoint(tg.astype(np.uint32), D, 5000, TOL)
    _, graw, gp = ref.gradients(D)
    assert np.array_equal(w.grad.numpy().astype(np.float32), graw)
    assert np.array_equal(pt.grad.numpy(), gp)


def test_autograd_quadratic_loss_matches_fp64_reference(N, O):
    g, og, wraw, p = problem(N, O, seed=7)
    a = np.random.default_rng(3).random(og.nV)
    pr = N.PageRank(g, N.F64, 1, weighted=True)
    w = torch.tensor(wraw.astype(np.float64), requires_grad=True)
    pt = torch.tensor(p, requires_grad=True)
    x = PPRFunction.apply(w, pt, pr)
    loss = (torch.from_numpy(a) * x * x).sum()
    loss.backward()
    xr, graw_ref, gp_ref = oracle_grads(O, og, wraw, p, lambda xx: 2.0 * a * xx)
    assert rel_l1(x.detach().numpy(), xr) < 1e-11
    assert rel_l1(w.grad.numpy(), graw_ref) < 1e-6
    assert rel_l1(pt.grad.numpy(), gp_ref) < 1e-9


def test_autograd_central_differences(N, O):
    g, og, wraw, p = problem(N, O, seed=9)
    a = np.random.default_rng(4).random(og.nV)

    def loss_of(w_np, p_np, pr):
        x = PPRFunction.apply(torch.from_numpy(w_np), torch.from_numpy(p_np), pr)
        return float((torch.from_numpy(a) * x * x).sum())

    pr = N.PageRank(g, N.F64, 1, weighted=True)
    w = torch.tensor(wraw.astype(np.float64), requires_grad=True)
    pt = torch.tensor(p, requires_grad=True)
    x = PPRFunction.apply(w, pt, pr)
    (torch.from_numpy(a) * x * x).sum().backward()
    gw = w.grad.numpy()
    gpv = pt.grad.numpy()
    fd_pr = N.PageRank(g, N.F64, 1, weighted=True)
    rng = np.random.default_rng(5)
    for e in rng.choice(og.nE, 8, replace=False):
        h = 1e-3 * float(wraw[e])
        w_hi = wraw.astype(np.float64).copy()
        w_lo = w_hi.copy()
        w_hi[e] = np.float32(wraw[e] + h)
        w_lo[e] = np.float32(wraw[e] - h)
        fd = (loss_of(w_hi, p, fd_pr) - loss_of(w_lo, p, fd_pr)) / (w_hi[e] - w_lo[e])
        assert abs(fd - gw[e]) <= 2e-4 * max(abs(fd), 1e-9) + 1e-9
    for v in rng.choice(og.nV, 8, replace=False):
        h = 1e-3
        p_hi = p.copy()
        p_lo = p.copy()
        p_hi[v] += h
        p_lo[v] -= h
        fd = (loss_of(wraw.astype(np.float64), p_hi, fd_pr) - loss_of(wraw.astype(np.float64), p_lo, fd_pr)) / (2 * h)
        assert abs(fd - gpv[v]) <= 2e-4 * max(abs(fd), 1e-9) + 1e-9


def test_autograd_batched_equals_per_vector(N, O):
    k = 4
    g, og, wraw, P = problem(N, O, seed=6, k=k)
    A = np.random.default_rng(8).random((og.nV, k))
    pr = N.PageRank(g, N.F64, k, weighted=True)
    w = torch.tensor(wraw.astype(np.float64), requires_grad=True)
    pt = torch.tensor(P, requires_grad=True)
    X = PPRFunction.apply(w, pt, pr)
    (torch.from_numpy(A) * X * X).sum().backward()
    graw_sum = np.zeros(og.nE)
    for j in range(k):
        xr, graw_j, gp_j = oracle_grads(O, og, wraw, P[:, j], lambda xx, j=j: 2.0 * A[:, j] * xx)
        assert rel_l1(X.detach().numpy()[:, j], xr) < 1e-11
        assert rel_l1(pt.grad.numpy()[:, j], gp_j) < 1e-9
        graw_sum += graw_j
    assert rel_l1(w.grad.numpy(), graw_sum) < 1e-6
