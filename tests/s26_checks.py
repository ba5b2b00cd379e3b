"""Optimality checks of BFS levels, CC labels and SSSP distances on a large
graph, evaluated with torch on the GPU in a process of its own (torch's HIP
runtime and the library's do not share one process).  Used by
test_gpu_s26_invariants.py:

    python tests/s26_checks.py DIR

DIR holds off.npy (uint64 [nV+1]) and idx.npy (uint32 [nE]) of a symmetric CSR,
w.npy (uint32 [nE], CSR order), meta.npy ([sssp source, bfs sources...]),
lev.npy (int32 [nsrc, nV], one row per bfs source), lab.npy (uint64 [nV]) and
dist.npy (uint32 [nV], from meta[0], which is also the first bfs source).  Prints one "ok NAME" line per
passed check and exits non-zero at the first failure.
"""
import os
import sys

import numpy as np
import torch

BIG = 1 << 40


def load(d, name):
    return np.load(os.path.join(d, name + ".npy"), mmap_mode="r")


def main(d):
    dev = torch.device("cuda")
    off = torch.from_numpy(np.ascontiguousarray(load(d, "off")).view(np.int64)).to(dev)
    nV = off.numel() - 1
    dst = torch.from_numpy(np.ascontiguousarray(load(d, "idx")).view(np.int32)).to(dev).long()
    deg = off[1:] - off[:-1]
    src = torch.repeat_interleave(torch.arange(nV, device=dev), deg)
    meta = np.asarray(load(d, "meta"))
    s = int(meta[0])

    # BFS levels, per source: level[src] = 0; |level[u] - level[v]| <= 1 on every
    # edge; the reached set is closed under edges; every other reached vertex
    # has a neighbour one level up
    levs = load(d, "lev")
    for i, si in enumerate(meta[1:]):
        si = int(si)
        L = torch.from_numpy(np.ascontiguousarray(levs[i])).to(dev).long()
        assert int(L[si]) == 0, f"bfs {si}: source level"
        lu, lv = L[src], L[dst]
        ru = lu >= 0
        assert torch.equal(ru, lv >= 0), f"bfs {si}: reached set not closed"
        assert int((lu[ru] - lv[ru]).abs().max()) <= 1, f"bfs {si}: level jump across an edge"
        up = torch.full((nV,), BIG, dtype=torch.long, device=dev)
        up.scatter_reduce_(0, dst, torch.where(ru, lu, torch.full_like(lu, BIG)), "amin")
        reached = L >= 0
        reached[si] = False
        assert torch.equal(up[reached], L[reached] - 1), f"bfs {si}: vertex without a parent one level up"
        print("ok bfs", si, int((L >= 0).sum()), "reached,", int(L.max()), "levels", flush=True)
        del lu, lv, up, ru
        if i == 0:
            L0 = L
    L = L0

    # CC labels: constant along every edge, label[v] <= v, label[label[v]] =
    # label[v]; and equal to an independent min-label propagation (each vertex
    # takes the smallest label among itself and its neighbours, then jumps to its
    # label's label, until nothing changes), whose fixpoint is the smallest
    # vertex id of each component: those checks alone would also pass two
    # components sharing one root
    lab = torch.from_numpy(np.ascontiguousarray(load(d, "lab")).view(np.int64)).to(dev)
    ar = torch.arange(nV, device=dev)
    assert torch.equal(lab[src], lab[dst]), "cc: label differs across an edge"
    assert bool((lab <= ar).all()), "cc: label above vertex id"
    assert torch.equal(lab[lab], lab), "cc: label is not its own label"
    M = ar.clone()
    for rounds in range(1, 200):
        nxt = M.clone()
        nxt.scatter_reduce_(0, dst, M[src], "amin")
        nxt = nxt[nxt]
        if torch.equal(nxt, M):
            break
        M = nxt
    assert torch.equal(M, lab), "cc: labels differ from min-label propagation"
    print("ok cc", int((lab == ar).sum()), "components,", rounds, "propagation rounds", flush=True)
    del lab, M, nxt

    # SSSP: dist[s] = 0; nothing to relax; every other reached vertex has a
    # tight in-edge; reached set = BFS reached set
    w = torch.from_numpy(np.ascontiguousarray(load(d, "w")).view(np.int32)).to(dev).long()
    D = torch.from_numpy(np.ascontiguousarray(load(d, "dist")).astype(np.int64)).to(dev)
    inf = 0xFFFFFFFF
    assert int(D[s]) == 0, "sssp: source distance"
    du = D[src]
    rd = du != inf
    assert bool((D[dst][rd] <= du[rd] + w[rd]).all()), "sssp: relaxable edge"
    best = torch.full((nV,), BIG, dtype=torch.long, device=dev)
    best.scatter_reduce_(0, dst, torch.where(rd, du + w, torch.full_like(du, BIG)), "amin")
    r2 = D != inf
    r2[s] = False
    assert torch.equal(best[r2], D[r2]), "sssp: vertex without a tight in-edge"
    assert torch.equal(D != inf, L >= 0), "sssp: reached set differs from bfs"
    print("ok sssp", int(D[D != inf].max()), "max distance", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1]))
