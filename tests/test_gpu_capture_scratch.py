"""A captured fixed-iteration PageRank forward replayed after another handle
has grown (and so moved) the library's shared reduction scratch.

The single-vector forward's init pass reduces the dangling mass through that
scratch; a batched handle asks for a larger one, which frees the old block.
The replay must notice the move and capture again rather than write to freed
memory.  The batched size here (k = 8 at scale 14) exceeds what the tests
that run before this file request, so the scratch is regrown in this test
whether it runs alone or in the full suite.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_replay_after_batched_run_regrows_scratch(N, dtype):
    dt = N.F64 if dtype == "f64" else N.F32
    g = N.Graph.rmat(14, 16, 8, False)
    eager = N.PageRank(g, dt)
    eager.set_graphs(False)
    eager.run(0.85, 20)
    ref = eager.ranks()
    a = N.PageRank(g, dt)
    a.run(0.85, 20)                       # captured, then replayed below
    assert np.array_equal(a.ranks(), ref)
    k = 8
    rng = np.random.default_rng(8)
    P = np.zeros((g.nV, k), np.float64 if dtype == "f64" else np.float32)
    for j in range(k):
        P[rng.choice(g.nV, 30, replace=False), j] = 1.0 / 30
    b = N.PageRank(g, dt, k, weighted=True)
    b.set_edge_weights(None, seed=8)
    b.set_personalization(P)
    b.run(0.85, 10)
    xb = b.ranks()
    assert np.all(np.abs(xb.sum(axis=0, dtype=np.float64) - 1.0) < 1e-5)
    for _ in range(3):
        a.run(0.85, 20)
        assert np.array_equal(a.ranks(), ref)
        b.run(0.85, 10)
        assert np.array_equal(b.ranks(), xb)
    b.close()
    a.run(0.85, 20)
    assert np.array_equal(a.ranks(), ref)
