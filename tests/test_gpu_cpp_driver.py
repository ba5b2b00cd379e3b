"""GPU end-to-end: GRE's text graph formats -> DistributedGraph::ingress ->
SynchronousEngine<Program>::run<dir> on the device, through the C++ host layer
(graphine_amd/include) and the C ABI, against the oracle's GRE engine.

The driver binary (tests/cpp/_build/gre_driver) is built on the CPU by
__graft_entry__.build(); it follows the call sequence of GRE's usr/test_*.cpp.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "_build", "gre_driver")


def write_graph(tmp, nV, src, dst, roots=None, weights=None, files=1):
    prefix = str(tmp / "g")
    with open(prefix + ".meta", "w") as f:
        f.write(f"{nV} {len(src)}\n")
    # split across several .edges.i files, in order (GraphIngress reads i = 0, 1, ...)
    parts = np.array_split(np.arange(len(src)), files)
    for i, part in enumerate(parts):
        with open(f"{prefix}.edges.{i}", "w") as f:
            f.write("# generated edge list\n")
            f.write("".join(f"{src[j]}\t{dst[j]}\n" for j in part))
    if roots is not None:
        with open(prefix + ".roots", "w") as f:
            f.write(f"{len(roots)}\n" + "".join(f"{r}\n" for r in roots))
    if weights is not None:
        with open(prefix + ".weights", "w") as f:
            f.write("\n".join(str(int(w)) for w in weights) + "\n")
    return prefix


def run_driver(prog, prefix, tmp):
    assert os.path.exists(DRIVER), "build first (__graft_entry__.build())"
    out = str(tmp / f"{prog}.bin")
    r = subprocess.run([DRIVER, prog, prefix, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    counts = [int(l.rsplit("...", 1)[1]) for l in r.stdout.splitlines() if l.startswith("Superstep-")]
    ss = int([l for l in r.stdout.splitlines() if l.startswith("supersteps")][-1].split()[1])
    return np.fromfile(out, np.uint8), counts, ss, r.stdout


def edges_of(og):
    src = np.repeat(np.arange(og.nV, dtype=np.uint32), og.outdeg)
    return src, og.idx.copy()


def test_cpp_pagerank_from_files(O, tmp_path):
    og = O.rmat_graph(12, 16, 1, False)
    s, d = edges_of(og)
    prefix = write_graph(tmp_path, og.nV, s, d, files=3)
    raw, counts, ss, _ = run_driver("pagerank", prefix, tmp_path)
    pr = raw.view(np.float64)
    st = O.default_state(O.PAGERANK, og.nV)
    oss, ocounts, _ = O.gre_run(O.PAGERANK, O.DIR_OUT, og, st, max_ss=10)
    assert ss == oss == 10 and counts == ocounts.tolist()
    assert np.allclose(pr, st[2], rtol=1e-12, atol=0)


def test_cpp_bfs_cc_from_shuffled_files(O, tmp_path):
    og = O.rmat_graph(12, 16, 2, True)
    s, d = edges_of(og)
    perm = np.random.default_rng(0).permutation(len(s))
    s, d = s[perm], d[perm]
    g_in = O.graph_from_edges(og.nV, s, d)          # GRE's input-order CSR/CSC
    root = int(np.argmax(og.outdeg))
    prefix = write_graph(tmp_path, og.nV, s, d, roots=[root, 5, 7])
    raw, counts, ss, _ = run_driver("bfs", prefix, tmp_path)
    st = O.default_state(O.BFS, og.nV)
    oss, ocounts, _ = O.gre_run(O.BFS, O.DIR_IN, g_in, st, srcs=[root])
    assert ss == oss and counts == ocounts.tolist()
    assert np.array_equal(raw.view(np.uint64), st[2])
    raw, counts, ss, out = run_driver("cc", prefix, tmp_path)
    st = O.default_state(O.CC, og.nV)
    oss, ocounts, _ = O.gre_run(O.CC, O.DIR_INOUT, g_in, st)
    lab = raw.view(np.uint64)
    assert ss == oss and counts == ocounts.tolist() and np.array_equal(lab, st[2])
    # the program's reduce() ran on host labels every superstep
    ncomp = int(out.split("components ")[1].split()[0])
    nred = int(out.split("reduces ")[1].split()[0])
    assert ncomp == len(np.unique(lab)) and nred == ss + 1


def test_cpp_sssp_with_edge_property_file(O, tmp_path):
    og = O.rmat_graph(12, 16, 2, True)
    s, d = edges_of(og)
    w = O.sssp_weights(og, seed=3)                  # local CSR order, as loadEdgeProperty reads
    root = int(np.argmax(og.outdeg))
    prefix = write_graph(tmp_path, og.nV, s, d, roots=[root, root], weights=w)
    raw, counts, ss, _ = run_driver("sssp", prefix, tmp_path)
    st = O.default_state(O.SSSP, og.nV)
    oss, ocounts, _ = O.gre_run(O.SSSP, O.DIR_OUT, og, st, aux=w, srcs=[root], max_ss=128)
    assert ss == oss and counts == ocounts.tolist()
    assert np.array_equal(raw.view(np.uint32), st[1])
    assert np.array_equal(raw.view(np.uint32), O.dijkstra(og, w, root))


def test_cpp_bc_two_engine_pipeline(O, tmp_path):
    og = O.rmat_graph(11, 16, 4, True)
    s, d = edges_of(og)
    root = int(np.argmax(og.outdeg))
    prefix = write_graph(tmp_path, og.nV, s, d, roots=[root, root])
    raw, counts, ss, _ = run_driver("bc", prefix, tmp_path)
    n = og.nV
    vtx = raw[: 8 * n].view(O.VTX_DTYPE)
    deps = raw[8 * n:].view(np.float64)
    fst = O.default_state(O.BC_FORWARD, n)
    oss, ocounts, level = O.gre_run(O.BC_FORWARD, O.DIR_OUT, og, fst, srcs=[root], level=1)
    assert np.array_equal(vtx, fst[2])
    bst = O.default_state(O.BC_BACKWARD, n)
    for i in range(level - 2, -1, -1):
        O.gre_run(O.BC_BACKWARD, O.DIR_IN, og, bst, aux=fst[2], level=i)
    assert np.allclose(deps, bst[2], rtol=1e-12, atol=1e-12)
    sig, dl, lv = O.brandes_one(og, root)
    assert np.allclose(deps[lv >= 0], dl[lv >= 0], rtol=1e-9, atol=1e-12)


def test_cpp_countin(O, tmp_path):
    og = O.rmat_graph(12, 8, 3, False)
    s, d = edges_of(og)
    prefix = write_graph(tmp_path, og.nV, s, d)
    raw, counts, ss, _ = run_driver("countin", prefix, tmp_path)
    cn = raw.view(np.int32)
    assert cn.sum() == og.nE and np.array_equal(cn, og.outdeg)   # run<In>: one message per out-edge


def run_driver_rc(prog, prefix, tmp):
    out = str(tmp / f"{prog}.bin")
    return subprocess.run([DRIVER, prog, prefix, out], capture_output=True, text=True, timeout=120)


def test_cpp_binding_check_rejects_a_program_with_other_rules(O, tmp_path):
    # PageRank02 is bound to PageRank's device program but teleports 0.2: the
    # check on the probe graph catches it before anything runs on the input
    og = O.rmat_graph(10, 8, 1, False)
    s, d = edges_of(og)
    prefix = write_graph(tmp_path, og.nV, s, d)
    r = run_driver_rc("pagerank02", prefix, tmp_path)
    assert r.returncode == 5, r.stdout + r.stderr
    assert "does not compute the rules of device program 0" in r.stderr
    assert "Superstep-" not in r.stdout


def test_cpp_unbound_program_is_matched_by_its_rules(O, tmp_path):
    # Reach has no ProgramTraits binding; it computes BFS's rules in its own way
    # and runs on the BFS device program, with GRE's BFS results
    og = O.rmat_graph(12, 16, 2, True)
    s, d = edges_of(og)
    root = int(np.argmax(og.outdeg))
    prefix = write_graph(tmp_path, og.nV, s, d, roots=[root, root])
    raw, counts, ss, out = run_driver("reach", prefix, tmp_path)
    assert "device program 2" in out
    st = O.default_state(O.BFS, og.nV)
    oss, ocounts, _ = O.gre_run(O.BFS, O.DIR_IN, og, st, srcs=[root])
    assert ss == oss and counts == ocounts.tolist()
    assert np.array_equal(raw.view(np.uint64), st[2])


def test_cpp_unbound_program_without_a_match_is_refused(O, tmp_path):
    # MaxLabel, run over every vertex, labels each component with its largest
    # vertex id; no device program computes that
    og = O.rmat_graph(10, 8, 2, True)
    s, d = edges_of(og)
    prefix = write_graph(tmp_path, og.nV, s, d, roots=[0, 0])
    r = run_driver_rc("maxlabel", prefix, tmp_path)
    assert r.returncode == 5, r.stdout + r.stderr
    assert "matches no device program" in r.stderr
