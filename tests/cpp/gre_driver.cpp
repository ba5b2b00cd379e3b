// End-to-end driver for the GRE-compatible C++ host layer (GPU tests).
//
// Follows the call sequence of GRE's drivers (usr/test_*.cpp in the GRE tree):
// ingress from <prefix>.meta / <prefix>.edges.* / <prefix>.roots, then
// SynchronousEngine<Program>::initialize / run<dir> / finalize, and writes
// the program's result arrays to <out> as raw bytes for the Python test to
// compare against the oracle.
//
// The vertex programs below are this test's own: minimal classes with the
// names of GRE's shipped programs, so they bind to the same device kernels.
// They declare only what the engine needs (alloc/free and, for BC, the level
// accessors); GRE's own program headers compile against the same engine and
// are checked by tests/test_dropin_compile.py where the GRE tree is present.
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

class PageRank : public GRE::VertexCode<double, double, double, DG> {
 public:
  explicit PageRank(DG& g) : GRE::VertexCode<double, double, double, DG>(g), pr(NULL), ps(NULL), pc(NULL) {}
  void alloc() { allocScatterData(ps); allocCombineData(pc, 0.0); allocVertexData(pr); }
  void free() { freeScatterData(ps); freeCombineData(pc); freeVertexData(pr); }
  double *pr, *ps, *pc;
};

class BFS : public GRE::VertexCode<vid_t, vid_t, vid_t, DG> {
 public:
  explicit BFS(DG& g) : GRE::VertexCode<vid_t, vid_t, vid_t, DG>(g), pred(NULL), ps(NULL), pc(NULL) {}
  void alloc() { allocScatterData(ps); allocCombineData(pc, (vid_t)-1); allocVertexData(pred, (vid_t)-1); }
  void free() { freeScatterData(ps); freeCombineData(pc); freeVertexData(pred); }
  vid_t *pred, *ps, *pc;
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
  uint32_t *states, *last, *w;
};

class Forward : public GRE::VertexCode<int, int, vtxData, DG> {
 public:
  Forward(DG& g, vtxData*& st) : GRE::VertexCode<int, int, vtxData, DG>(g), state(st), np(NULL), sum(NULL), currLevel(1) {}
  void alloc() { allocScatterData(np, 0); allocCombineData(sum, 0); allocVertexData(state); }
  void free() { freeScatterData(np); freeCombineData(sum); }
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
  Backward(DG& g, vtxData* bs, double*& dep)
      : GRE::VertexCode<double, double, double, DG>(g), bfsState(bs), deps(dep), pw(NULL), sum(NULL), currLevel(0) {}
  void alloc() { allocScatterData(pw, 0); allocCombineData(sum, 0); allocVertexData(deps, 0.0); }
  void free() { freeScatterData(pw); freeCombineData(sum); }
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
  const bool undirected = prog != "pagerank" && prog != "countin";
  bool ok = undirected ? g.ingress<InOut>(comm, pool, prefix, "ascii", "edgelists")
                       : g.ingress(comm, pool, prefix, "ascii", "edgelists");
  if (!ok) return 3;
  const size_t n = g.get_num_masters();
  std::vector<vid_t> src;
  SrcSet srcSet;
  std::ofstream o(out.c_str(), std::ios::binary);
  long ss = -1;
  try {
    if (prog == "pagerank") {
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
      if (prog == "bfs") {
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
