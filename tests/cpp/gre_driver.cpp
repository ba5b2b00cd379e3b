// End-to-end driver for the GRE-compatible C++ host layer (GPU tests).
//
// Follows the call sequence of GRE's drivers (usr/test_*.cpp in the GRE tree):
// ingress from <prefix>.meta / <prefix>.edges.* / <prefix>.roots, then
// SynchronousEngine<Program>::initialize / run<dir> / finalize, and writes
// the program's result arrays to <out> as raw bytes for the Python test to
// compare against the oracle.
//
// The vertex programs below are this test's own implementations of the rules
// of GRE's shipped programs (include/Kernels/*.hpp in the GRE tree), under the
// same class names, so they bind to the same device programs.  At each class's
// first run<dir>() the engine checks, on a small probe graph, that these
// methods compute what the bound device program computes.  Three more classes
// exercise that check: PageRank02 is bound to PageRank's device program but
// uses damping 0.2 (rejected), Reach has no binding and computes BFS's rules in
// its own way (matched to the BFS device program), MaxLabel has no binding and
// matches no device program (rejected).
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "GRE.hpp"

typedef GRE::Graph::DistributedGraph<GRE::Graph::GraphInCSR> DG;
typedef GRE::Graph::SourceVertexSet SrcSet;

struct vtxData {
  int level;
  int numPaths;
  vtxData() : level(-1), numPaths(0) {}
};

static const vid_t kNone = (vid_t)-1;

class PageRank : public GRE::VertexCode<double, double, double, DG> {
 public:
  explicit PageRank(DG& g, double d = 0.15) : GRE::VertexCode<double, double, double, DG>(g), pr(NULL), ps(NULL), pc(NULL), teleport(d) {}
  void alloc() { allocScatterData(ps); allocCombineData(pc, 0.0); allocVertexData(pr); }
  void free() { freeScatterData(ps); freeCombineData(pc); freeVertexData(pr); }
  bool init(const lvid_t v) {
    pr[v] = teleport;
    const size_t d = graph.getOutdegree(v);
    if (d == 0) return false;
    ps[v] = pr[v] / (double)d;
    return true;
  }
  template <typename Ctx, typename It>
  void scatter(Ctx& ctx, const lvid_t u, const lvid_t v, const It&) {
    double m = ps[u];
    ctx.engine->sendMessage(ctx, v, m);
  }
  template <typename Lock>
  bool combine(const lvid_t v, const double& m, Lock& lk) {
    lk.spinLock();
    pc[v] += m;
    lk.unLock();
    return true;
  }
  bool apply(const lvid_t v) {
    pr[v] = teleport + (1.0 - teleport) * pc[v];
    pc[v] = 0.0;
    return false;
  }
  bool assert_to_halt(const lvid_t) { return false; }
  double *pr, *ps, *pc;
  double teleport;
};

// PageRank's device program, bound explicitly, but a teleport of 0.2
class PageRank02 : public PageRank {
 public:
  explicit PageRank02(DG& g) : PageRank(g, 0.2) {}
};
namespace GRE {
namespace Device {
template <>
struct ProgramTraits<PageRank02> : ProgramTraits< ::PageRank> {};
}  // namespace Device
}  // namespace GRE

class BFS : public GRE::VertexCode<vid_t, vid_t, vid_t, DG> {
 public:
  explicit BFS(DG& g) : GRE::VertexCode<vid_t, vid_t, vid_t, DG>(g), pred(NULL), ps(NULL), pc(NULL) {}
  void alloc() { allocScatterData(ps); allocCombineData(pc, kNone); allocVertexData(pred, kNone); }
  void free() { freeScatterData(ps); freeCombineData(pc); freeVertexData(pred); }
  bool init(const lvid_t v) {
    pred[v] = graph.get_vid(v);
    ps[v] = pred[v];
    return true;
  }
  template <typename Ctx, typename It>
  void scatter(Ctx& ctx, const lvid_t u, const lvid_t v, const It&) {
    vid_t m = ps[u];
    ctx.engine->sendMessage(ctx, v, m);
  }
  template <typename Lock>
  bool combine(const lvid_t v, const vid_t& m, Lock& lk) {
    if (pc[v] != kNone) return false;
    lk.spinLock();
    const bool won = pc[v] == kNone;
    if (won) pc[v] = m;
    lk.unLock();
    return won;
  }
  bool apply(const lvid_t v) {
    if (pred[v] != kNone) return false;
    pred[v] = pc[v];
    ps[v] = pred[v];
    return true;
  }
  vid_t *pred, *ps, *pc;
};

// BFS's rules with its own data layout and no binding: the root of each
// reached vertex, first message wins
class Reach : public GRE::VertexCode<vid_t, vid_t, vid_t, DG> {
 public:
  explicit Reach(DG& g) : GRE::VertexCode<vid_t, vid_t, vid_t, DG>(g), root(NULL), out(NULL), in(NULL) {}
  void alloc() { allocScatterData(out); allocCombineData(in, kNone); allocVertexData(root, kNone); }
  void free() { freeScatterData(out); freeCombineData(in); freeVertexData(root); }
  bool init(const lvid_t v) { out[v] = root[v] = (vid_t)v; return true; }
  template <typename Ctx, typename It>
  void scatter(Ctx& ctx, const lvid_t u, const lvid_t v, const It&) {
    vid_t m = out[u];
    ctx.engine->sendMessage(ctx, v, m);
  }
  template <typename Lock>
  bool combine(const lvid_t v, const vid_t& m, Lock& lk) {
    lk.spinLock();
    const vid_t was = in[v];
    if (was == kNone) in[v] = m;
    lk.unLock();
    return was == kNone;
  }
  bool apply(const lvid_t v) {
    const bool fresh = root[v] == kNone;
    if (fresh) out[v] = root[v] = in[v];
    return fresh;
  }
  vid_t *root, *out, *in;
};

// the largest label of each component: no device program computes this
class MaxLabel : public GRE::VertexCode<vid_t, vid_t, vid_t, DG> {
 public:
  explicit MaxLabel(DG& g) : GRE::VertexCode<vid_t, vid_t, vid_t, DG>(g), lab(NULL), ls(NULL), lc(NULL) {}
  void alloc() { allocScatterData(ls); allocCombineData(lc, 0); allocVertexData(lab, 0); }
  void free() { freeScatterData(ls); freeCombineData(lc); freeVertexData(lab); }
  bool init(const lvid_t v) { lab[v] = ls[v] = lc[v] = (vid_t)v; return true; }
  template <typename Ctx, typename It>
  void scatter(Ctx& ctx, const lvid_t u, const lvid_t v, const It&) {
    vid_t m = ls[u];
    ctx.engine->sendMessage(ctx, v, m);
  }
  template <typename Lock>
  bool combine(const lvid_t v, const vid_t& m, Lock& lk) {
    lk.spinLock();
    const bool up = m > lc[v];
    if (up) lc[v] = m;
    lk.unLock();
    return up;
  }
  bool apply(const lvid_t v) {
    if (lc[v] <= lab[v]) return false;
    ls[v] = lab[v] = lc[v];
    return true;
  }
  vid_t *lab, *ls, *lc;
};

class ConnectedComponents : public GRE::VertexCode<vid_t, vid_t, vid_t, DG> {
 public:
  explicit ConnectedComponents(DG& g)
      : GRE::VertexCode<vid_t, vid_t, vid_t, DG>(g), cc(NULL), cs(NULL), cm(NULL), reduces(0), lastCount(0) {}
  void alloc() {
    allocScatterData(cs);
    allocCombineData(cm, std::numeric_limits<vid_t>::max());
    allocVertexData(cc, std::numeric_limits<vid_t>::max());
  }
  void free() { freeScatterData(cs); freeCombineData(cm); freeVertexData(cc); }
  bool init(const lvid_t v) {
    cc[v] = graph.get_vid(v);
    cs[v] = cm[v] = cc[v];
    return true;
  }
  template <typename Ctx, typename It>
  void scatter(Ctx& ctx, const lvid_t u, const lvid_t v, const It&) {
    vid_t m = cs[u];
    ctx.engine->sendMessage(ctx, v, m);
  }
  template <typename Lock>
  bool combine(const lvid_t v, const vid_t& m, Lock& lk) {
    if (!(m < cm[v])) return false;
    lk.spinLock();
    const bool lower = m < cm[v];
    if (lower) cm[v] = m;
    lk.unLock();
    return lower;
  }
  bool apply(const lvid_t v) {
    if (!(cm[v] < cc[v])) return false;
    cs[v] = cc[v] = cm[v];
    return true;
  }
  // GRE's CC counts components in reduce() from the host labels every superstep
  template <typename DC>
  void reduce(DC& dc) {
    long n = 0;
    for (size_t v = 0; v < graph.get_num_masters(); ++v) n += (cc[v] == graph.get_vid((lvid_t)v));
    long all = 0;
    dc.refUnderlying().reduce(&n, &all, 1, MPI_LONG, MPI_SUM, 0);
    lastCount = all;
    ++reduces;
  }
  vid_t *cc, *cs, *cm;
  long reduces, lastCount;
};

class SSSP : public GRE::VertexCode<uint32_t, uint32_t, uint32_t, DG> {
 public:
  explicit SSSP(DG& g) : GRE::VertexCode<uint32_t, uint32_t, uint32_t, DG>(g), states(NULL), last(NULL), w(NULL) {}
  void alloc() {
    allocScatterData(last);
    allocCombineData(states, std::numeric_limits<uint32_t>::max());
    allocVertexData(states);
  }
  void free() {
    freeScatterData(last);
    freeCombineData(states);
    freeVertexData(states);
  }
  void loadEdgeProperty(const std::string& f) { graph.loadEdgeProperty(w, f); }
  void loadEdgeProperty() { graph.loadEdgeProperty(w); }
  bool init(const lvid_t v) {
    states[v] = 0;
    last[v] = 0;
    return true;
  }
  template <typename Ctx, typename It>
  void scatter(Ctx& ctx, const lvid_t u, const lvid_t v, const It& it) {
    uint32_t d = last[u] + w[graph.refLocalGraph().leid(it)];
    ctx.engine->sendMessage(ctx, v, d);
  }
  template <typename Lock>
  bool combine(const lvid_t v, const uint32_t& d, Lock& lk) {
    if (!(d < states[v])) return false;
    lk.spinLock();
    const bool shorter = d < states[v];
    if (shorter) states[v] = d;
    lk.unLock();
    return shorter;
  }
  bool apply(const lvid_t v) {
    last[v] = states[v];
    return true;
  }
  uint32_t *states, *last, *w;
};

class Forward : public GRE::VertexCode<int, int, vtxData, DG> {
 public:
  Forward(DG& g, vtxData*& st, int level = 1)
      : GRE::VertexCode<int, int, vtxData, DG>(g), state(st), np(NULL), sum(NULL), currLevel(level) {}
  void alloc() { allocScatterData(np, 0); allocCombineData(sum, 0); allocVertexData(state); }
  void free() { freeScatterData(np); freeCombineData(sum); }
  bool init(const lvid_t v) {
    state[v].level = 0;
    state[v].numPaths = 1;
    np[v] = 1;
    return true;
  }
  template <typename Ctx, typename It>
  void scatter(Ctx& ctx, const lvid_t u, const lvid_t v, const It&) {
    int m = np[u];
    ctx.engine->sendMessage(ctx, v, m);
  }
  template <typename Lock>
  bool combine(const lvid_t v, const int& m, Lock& lk) {
    lk.spinLock();
    sum[v] = (int)((unsigned)sum[v] + (unsigned)m);   // GRE's int sum, wrapping
    lk.unLock();
    return true;
  }
  bool apply(const lvid_t v) {
    if (state[v].level >= 0) return false;
    state[v].level = currLevel;
    state[v].numPaths = sum[v];
    np[v] = sum[v];
    return true;
  }
  template <typename DC>
  void reduce(DC&) { currLevel++; }
  int getCurrLevel() { return currLevel; }
  void setCurrLevel(int l) { currLevel = l; }
  vtxData*& state;
  int *np, *sum;
  int currLevel;
};

class Backward : public GRE::VertexCode<double, double, double, DG> {
 public:
  Backward(DG& g, vtxData* bs, double*& dep, int level = 0)
      : GRE::VertexCode<double, double, double, DG>(g), bfsState(bs), deps(dep), pw(NULL), sum(NULL), currLevel(level) {}
  void alloc() { allocScatterData(pw, 0); allocCombineData(sum, 0); allocVertexData(deps, 0.0); }
  void free() { freeScatterData(pw); freeCombineData(sum); }
  bool init(const lvid_t v) {
    if (bfsState[v].level != currLevel) return false;
    pw[v] = (1.0 + deps[v]) / (double)bfsState[v].numPaths;
    return true;
  }
  template <typename Ctx, typename It>
  void scatter(Ctx& ctx, const lvid_t u, const lvid_t v, const It&) {
    double m = pw[u];
    ctx.engine->sendMessage(ctx, v, m);
  }
  template <typename Lock>
  bool combine(const lvid_t v, const double& m, Lock& lk) {
    lk.spinLock();
    sum[v] += m;
    lk.unLock();
    return true;
  }
  bool apply(const lvid_t v) {
    if (bfsState[v].level == currLevel - 1) deps[v] = sum[v] * (double)bfsState[v].numPaths;
    return false;
  }
  int getCurrLevel() { return currLevel; }
  void setCurrLevel(int l) { currLevel = l; }
  vtxData* bfsState;
  double*& deps;
  double *pw, *sum;
  int currLevel;
};

class CountIn : public GRE::VertexCode<int, int, int, DG> {
 public:
  explicit CountIn(DG& g) : GRE::VertexCode<int, int, int, DG>(g), cn(NULL), cs(NULL), cm(NULL) {}
  void alloc() { allocScatterData(cs); allocCombineData(cm, 0); allocVertexData(cn); }
  void free() { freeScatterData(cs); freeCombineData(cm); freeVertexData(cn); }
  bool init(const lvid_t v) {
    cn[v] = 0;
    cs[v] = 1;
    return true;
  }
  template <typename Ctx, typename It>
  void scatter(Ctx& ctx, const lvid_t u, const lvid_t v, const It&) {
    int m = cs[u];
    ctx.engine->sendMessage(ctx, v, m);
  }
  template <typename Lock>
  bool combine(const lvid_t v, const int& m, Lock& lk) {
    lk.spinLock();
    cm[v] += m;
    lk.unLock();
    return true;
  }
  bool apply(const lvid_t v) {
    cn[v] = cm[v];
    return false;
  }
  int *cn, *cs, *cm;
};

static void dump(std::ofstream& o, const void* p, size_t bytes) { o.write(static_cast<const char*>(p), bytes); }

int main(int argc, char** argv) {
  if (argc < 4) {
    std::cerr << "usage: gre_driver <prog> <prefix> <out> [weights]" << std::endl;
    return 2;
  }
  const std::string prog(argv[1]), prefix(argv[2]), out(argv[3]);
  GRE::COMM::Underlying comm;
  comm.init(argc, argv);
  GRE::Util::BufferPool pool;
  pool.config(4096, 8192 * 256, true);
  DG g;
  const bool undirected = prog != "pagerank" && prog != "pagerank02" && prog != "countin";
  bool ok = undirected ? g.ingress<InOut>(comm, pool, prefix, "ascii", "edgelists")
                       : g.ingress(comm, pool, prefix, "ascii", "edgelists");
  if (!ok) return 3;
  const size_t n = g.get_num_masters();
  std::vector<vid_t> src;
  SrcSet srcSet;
  std::ofstream o(out.c_str(), std::ios::binary);
  long ss = -1;
  try {
    if (prog == "pagerank02") {
      srcSet.setALL();
      PageRank02 p(g);
      GRE::SynchronousEngine<PageRank02> e(g, comm, pool, p);
      e.initialize();
      e.configMaxSupersteps(10);
      e.run(srcSet, GRE::Par_ThreadPool);
      ss = e.getSuperstepCount();
      e.finalize();
    } else if (prog == "pagerank") {
      srcSet.setALL();
      PageRank p(g);
      GRE::SynchronousEngine<PageRank> e(g, comm, pool, p);
      e.initialize();
      e.configMaxSupersteps(10);
      e.run(srcSet, GRE::Par_ThreadPool);
      ss = e.getSuperstepCount();
      dump(o, p.pr, 8 * n);
      e.finalize();
    } else if (prog == "countin") {
      srcSet.setALL();
      CountIn c(g);
      GRE::SynchronousEngine<CountIn> e(g, comm, pool, c);
      e.initialize();
      e.configMaxSupersteps(1);
      e.run<In>(srcSet, GRE::Par_ThreadPool);
      ss = e.getSuperstepCount();
      dump(o, c.cn, 4 * n);
      e.finalize();
    } else {
      const size_t nr = GRE::Util::loadFromFile(prefix + ".roots", src, 0);
      if (nr == 0) return 4;
      if (prog == "reach" || prog == "maxlabel") {
        srcSet.addVertex(src[1], g);
        if (prog == "reach") {
          Reach r(g);
          GRE::SynchronousEngine<Reach> e(g, comm, pool, r);
          e.initialize();
          e.run<In>(srcSet, GRE::Par_ThreadPool);
          ss = e.getSuperstepCount();
          std::cout << "device program " << e.deviceProgram() << std::endl;
          dump(o, r.root, 8 * n);
          e.finalize();
        } else {
          srcSet.setALL();
          MaxLabel m(g);
          GRE::SynchronousEngine<MaxLabel> e(g, comm, pool, m);
          e.initialize();
          e.run<InOut>(srcSet, GRE::Par_ThreadPool);
          ss = e.getSuperstepCount();
          e.finalize();
        }
      } else if (prog == "bfs") {
        srcSet.addVertex(src[1], g);
        BFS b(g);
        GRE::SynchronousEngine<BFS> e(g, comm, pool, b);
        e.initialize();
        e.run<In>(srcSet, GRE::Par_ThreadPool);
        ss = e.getSuperstepCount();
        dump(o, b.pred, 8 * n);
        e.finalize();
      } else if (prog == "cc") {
        srcSet.setALL();
        ConnectedComponents c(g);
        GRE::SynchronousEngine<ConnectedComponents> e(g, comm, pool, c);
        e.initialize();
        e.run<InOut>(srcSet, GRE::Par_ThreadPool);
        ss = e.getSuperstepCount();
        dump(o, c.cc, 8 * n);
        std::cout << "components " << c.lastCount << " reduces " << c.reduces << std::endl;
        e.finalize();
      } else if (prog == "sssp") {
        srcSet.addVertex(src[1], g);
        SSSP s(g);
        s.loadEdgeProperty(argc > 4 ? argv[4] : prefix + ".weights");
        GRE::SynchronousEngine<SSSP> e(g, comm, pool, s);
        e.initialize();
        e.configMaxSupersteps(128);
        e.run(srcSet, GRE::Par_ThreadPool);
        ss = e.getSuperstepCount();
        dump(o, s.states, 4 * n);
        e.finalize();
      } else if (prog == "bc") {
        srcSet.addVertex(src[1], g);
        vtxData* st = NULL;
        Forward f(g, st);
        GRE::SynchronousEngine<Forward> fe(g, comm, pool, f);
        fe.initialize();
        fe.run<Out>(srcSet, GRE::Par_ThreadPool);
        ss = fe.getSuperstepCount();
        const int nLevels = f.getCurrLevel();
        fe.finalize();
        double* deps = NULL;
        Backward bw(g, st, deps);
        GRE::SynchronousEngine<Backward> be(g, comm, pool, bw);
        be.initialize();
        for (int i = nLevels - 2; i >= 0; --i) {
          srcSet.setALL();
          bw.setCurrLevel(i);
          be.run<In>(srcSet, GRE::Par_ThreadPool);
        }
        dump(o, st, 8 * n);
        dump(o, deps, 8 * n);
        be.finalize();
        delete[] st;
        delete[] deps;
      } else {
        std::cerr << "unknown program " << prog << std::endl;
        return 2;
      }
    }
  } catch (const std::exception& ex) {
    std::cerr << "error: " << ex.what() << std::endl;
    return 5;
  }
  std::cout << "supersteps " << ss << std::endl;
  comm.finalize();
  pool.finalize();
  return 0;
}
