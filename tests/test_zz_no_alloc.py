"""gre_pr_run_noalloc, the allocation-free form of gre_pr_run, called through
the C ABI: on a handle that one gre_pr_run has prepared it gives gre_pr_run's
ranks, residual and iteration count bit for bit, on two calls in a row (the
second on a smaller graph), with the library's device memory unchanged; it
refuses unprepared handles, batched handles and too many iterations."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_noalloc(N, pr, damping, iters, tol=0.0, want_residual=True):
    it, res = C.c_int(-1), C.c_double(-1.0)
    rc = N.lib().gre_pr_run_noalloc(pr.h, damping, iters, tol, C.byref(it),
                                    C.byref(res) if want_residual else None)
    return rc, it.value, res.value


def checked_noalloc(N, pr, damping, iters, tol=0.0, want_residual=True):
    """run_noalloc, plus the change of the library's device memory across it."""
    N.sync()
    before = N.mem_in_use()
    rc, it, res = run_noalloc(N, pr, damping, iters, tol, want_residual)
    N.sync()
    return rc, it, res, N.mem_in_use() - before


def prepared(N, g, dt, weighted=False):
    pr = N.PageRank(g, dt, 1, weighted=weighted)
    if weighted:
        pr.set_edge_weights(None, seed=4)
        rng = np.random.default_rng(4)
        p = np.zeros(g.nV)
        p[rng.choice(g.nV, 50, replace=False)] = 1.0 / 50
        pr.set_personalization(p)
    return pr


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_noalloc_equals_run_on_two_calls(N, dtype, weighted):
    dt = N.F64 if dtype == "f64" else N.F32
    for scale in (17, 12):                       # the second call gets a smaller input
        g = N.Graph.rmat(scale, 16, 9, False)
        pr = prepared(N, g, dt, weighted)
        it0, r0 = pr.run(0.85, 20)                # existing entry point (prepares the plan)
        x0 = pr.ranks()
        rc, it1, r1, grew = checked_noalloc(N, pr, 0.85, 20)
        assert rc == 0, N.lib().gre_last_error()
        assert grew == 0
        assert it1 == it0 == 20 and r1 == r0
        assert np.array_equal(pr.ranks(), x0)
        # converged runs (tol > 0) stop at the same iteration with the same bits
        tol = 1e-9 if dtype == "f64" else 1e-6
        it0, r0 = pr.run(0.85, 1000, tol)
        x0 = pr.ranks()
        rc, it1, r1, grew = checked_noalloc(N, pr, 0.85, 1000, tol)
        assert rc == 0 and grew == 0 and it1 == it0 and r1 == r0
        assert np.array_equal(pr.ranks(), x0)
        # enqueue-only form (no residual read-back), then the existing entry point again
        rc, it1, _, grew = checked_noalloc(N, pr, 0.85, 20, want_residual=False)
        assert rc == 0 and grew == 0 and it1 == 20
        x1 = pr.ranks()
        pr.run(0.85, 20)
        assert np.array_equal(pr.ranks(), x1)
        pr.close()
        g.close()


def test_noalloc_rejects_what_it_cannot_run(N):
    g = N.Graph.rmat(10, 16, 9, False)
    fresh = N.PageRank(g, N.F32)                # schedule not tuned, plan not built
    rc, _, _ = run_noalloc(N, fresh, 0.85, 20)
    assert rc == 2 and b"not built" in N.lib().gre_last_error()
    fresh.run(0.85, 5)
    rc, _, _ = run_noalloc(N, fresh, 0.85, 4097)
    assert rc == 2 and b"4096" in N.lib().gre_last_error()
    rc, _, _ = run_noalloc(N, fresh, 0.85, 4096, tol=1e-3)
    assert rc == 0
    batched = N.PageRank(g, N.F32, 4, weighted=True)
    rc, _, _ = run_noalloc(N, batched, 0.85, 20)
    assert rc == 8
    N.prof_enable(True)
    try:
        rc, _, _ = run_noalloc(N, fresh, 0.85, 20)
        assert rc == 1
    finally:
        N.prof_enable(False)
