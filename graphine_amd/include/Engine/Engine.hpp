// SynchronousEngine<VertexCode> (include/Engine/Engine.hpp:53-1418 in the GRE
// tree), running GRE's BSP superstep on one MI355X through the C ABI.
//
// The call sequence a GRE driver makes is unchanged:
//   SynchronousEngine<VC> engine(graph, comm, pool, vc);
//   engine.initialize();            // vc.alloc()
//   engine.configMaxSupersteps(n);
//   engine.run<dir>(srcSet, GRE::Par_ThreadPool);
//   engine.getSuperstepCount(); engine.get*Time();
//   engine.finalize();              // vc.free()
// run<dir>() uploads the program's scatter/combine/vertex arrays (whatever
// alloc() and the driver left in them), runs the init phase and the
// supersteps on the device (compute = push along dir, combine in HBM, apply,
// assert_to_halt, frontier rebuild), calls the program's reduce(dc) on the
// host after every superstep exactly where GRE does, prints GRE's
// "Superstep-k: Number of active vertices..." line, downloads the arrays back
// and calls output_statistics(dc).  The device kernels bound to the program in
// Engine/DevicePrograms.hpp compute its rules; the program's own
// scatter/combine/apply run only in the binding check: at the class's first
// run<dir>() in the process, both run on a small probe graph (the program's
// methods through GRE's superstep on the host, the bound kernels on the
// device) and must agree, else run<dir>() throws DeviceError.  A class with
// no binding is matched against the device programs the same way.
#ifndef GRE_AMD_ENGINE_HPP
#define GRE_AMD_ENGINE_HPP
#include <stdint.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <typeinfo>
#include <vector>

#include "Comm/BufferingComm.hpp"
#include "Comm/CommUnderlying.hpp"
#include "Config/UsrConfig.hpp"
#include "Engine/DevicePrograms.hpp"
#include "Graph/DistributedGraph.hpp"
#include "Graph/LocalGraph.hpp"
#include "Graph/VertexSet.hpp"
#include "Message.hpp"
#include "Parallel/ThreadContext.hpp"
#include "Status.hpp"
#include "Types.hpp"
#include "Util/BufferPool.hpp"
#include "Util/Util.hpp"
#include "Util/vLock.hpp"
#include "gre_dev.h"

namespace GRE {

typedef enum ParallelMode { Par_ThreadPool, Par_ThreadGroup, Par_Other } ParallelMode_t;

// A device call failed: the message carries gre_last_error().
struct DeviceError : std::runtime_error {
  int status;
  DeviceError(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void device_check(int rc, const char* what) {
  if (rc != 0) throw DeviceError(rc, std::string(what) + ": " + gre_last_error());
}

namespace Device {

// The probe graph of the binding check: 40 vertices, a directed graph with a
// cycle, a hub with in- and out-edges, a chain, vertices without out-edges,
// an isolated pair and two isolated vertices; edge properties loaded on it are
// 1 + (e * 7919) % 97 instead of random values.  Built once per process and
// kept until it ends.
template <class DG>
DG& probeGraph() {
  static DG* g = NULL;
  if (!g) {
    std::vector<uint32_t> src, dst;
    const size_t nV = 40;
    for (uint32_t v = 0; v < 20; ++v) { src.push_back(v); dst.push_back((v + 1) % 20); }    // cycle
    for (uint32_t v = 0; v < 20; v += 3) { src.push_back((v + 7) % 20); dst.push_back(v); }  // chords
    for (uint32_t v = 21; v < 32; ++v) { src.push_back(20); dst.push_back(v); }             // hub out
    for (uint32_t v = 24; v < 30; ++v) { src.push_back(v); dst.push_back(20); }             // hub in
    src.push_back(5); dst.push_back(20);
    src.push_back(22); dst.push_back(12);
    for (uint32_t v = 32; v < 35; ++v) { src.push_back(v); dst.push_back(v + 1); }          // chain 32..35
    src.push_back(9); dst.push_back(32);
    src.push_back(35); dst.push_back(17);
    src.push_back(36); dst.push_back(37);                                                    // isolated pair
    src.push_back(37); dst.push_back(36);
    std::vector<int> outdeg(nV, 0);
    for (size_t i = 0; i < src.size(); ++i) outdeg[src[i]]++;
    DG* p = new DG();
    p->setDeterministicProperties(true);
    if (!p->adoptEdges(nV, src, dst, outdeg)) {
      delete p;
      throw DeviceError(GRE_FAILURE, "building the probe graph of the binding check failed");
    }
    g = p;
  }
  return *g;
}

// Elementwise comparison of a host and a device result: integers and records
// byte for byte, floating point to 1e-12 relative (sums in another order).
template <class T>
inline bool sameElem(const T& a, const T& b, typename std::enable_if<std::is_floating_point<T>::value>::type* = 0) {
  if (std::memcmp(&a, &b, sizeof(T)) == 0) return true;
  if (!std::isfinite(a) || !std::isfinite(b)) return false;
  const T m = std::fabs(a) > std::fabs(b) ? std::fabs(a) : std::fabs(b);
  return std::fabs(a - b) <= (T)1e-12 * m;
}
template <class T>
inline bool sameElem(const T& a, const T& b, typename std::enable_if<!std::is_floating_point<T>::value>::type* = 0) {
  return std::memcmp(&a, &b, sizeof(T)) == 0;
}

// Silences std::cout while probe runs print GRE's per-superstep lines and
// statistics.
struct QuietCout {
  std::ostringstream sink;
  std::streambuf* old;
  QuietCout() : old(std::cout.rdbuf(sink.rdbuf())) {}
  ~QuietCout() { std::cout.rdbuf(old); }
};

}  // namespace Device

template <typename VertexCode>
class SynchronousEngine {
 public:
  typedef typename VertexCode::ScatterData_t ScatterData_t;
  typedef typename VertexCode::CombineData_t CombineData_t;
  typedef typename VertexCode::VertexData_t VertexData_t;
  typedef Graph::GraphInCSR LocalGraph_t;
  typedef Graph::DistributedGraph<LocalGraph_t> DistributedGraph_t;
  typedef Graph::SourceVertexSet SourceVertexSet_t;
  typedef Graph::VertexSet VertexSet_t;
  typedef COMM::BufComm<Util::BufferPool> BufComm_t;
  typedef SynchronousEngine<VertexCode> SELF_t;
  typedef HostContext<VertexSet_t, SELF_t> TP_Context_t;
  typedef Device::ProgramTraits<VertexCode> Traits;

  SynchronousEngine(DistributedGraph_t& g, COMM::Underlying& comm, Util::BufferPool& pool, VertexCode& vc)
      : graph_(g), vertexCode_(vc), dc_(comm, pool), dev_(NULL), isInitialized_(false), maxSupersteps_(1024),
        superstep_(0), weightsFrom_(NULL) {
    for (int i = 0; i < 5; ++i) times_[i] = 0.0;
    binding_.id = -1;
  }
  ~SynchronousEngine() {
    if (dev_) gre_engine_destroy(dev_);
  }

  void reset() {
    superstep_ = 0;
    vertexCode_.free();
    vertexCode_.alloc();
    for (int i = 0; i < 5; ++i) times_[i] = 0.0;
  }
  void configMaxSupersteps(const long n = 0) { maxSupersteps_ = n > 0 ? n : 1024; }
  void initialize() {
    if (!isInitialized_) {
      isInitialized_ = true;
      vertexCode_.alloc();
    }
  }
  void finalize() {
    if (isInitialized_) {
      isInitialized_ = false;
      vertexCode_.free();
    }
  }

  void run(SourceVertexSet_t& srcSet, const ParallelMode mode = Par_ThreadPool) { run<Out>(srcSet, mode); }

  template <Dir dir>
  void run(SourceVertexSet_t& srcSet, const ParallelMode mode = Par_ThreadPool) {
    static_assert(Traits::id < 0 || sizeof(ScatterData_t) == (size_t)Traits::scatterBytes,
                  "scatter data type does not match the binding");
    static_assert(Traits::id < 0 || sizeof(CombineData_t) == (size_t)Traits::combineBytes,
                  "combine data type does not match the binding");
    static_assert(Traits::id < 0 || sizeof(VertexData_t) == (size_t)Traits::vertexBytes,
                  "vertex data type does not match the binding");
    static_assert(dir == Out || dir == In || dir == InOut, "run<dir> needs Out, In or InOut");
    (void)mode;  // both GRE thread modes map to the same device superstep
    if (!isInitialized_) initialize();
    if (binding_.id < 0) binding_ = checkedBinding<dir>(srcSet.isALL());
    ensureDevice();
    device_check(gre_engine_config_max_supersteps(dev_, maxSupersteps_), "configMaxSupersteps");
    uploadState();
    if (binding_.usesLevel) device_check(gre_engine_set_level(dev_, levelOf(vertexCode_, 0)), "set level");
    lastCounts_.clear();

    const double t0 = now();
    const int gdir = dir == Out ? GRE_DIR_OUT : (dir == In ? GRE_DIR_IN : GRE_DIR_INOUT);
    if (srcSet.isALL()) {
      device_check(gre_engine_begin(dev_, gdir, NULL, 0), "init phase");
    } else {
      const std::vector<lvid_t>& v = srcSet.vertices;
      device_check(gre_engine_begin(dev_, gdir, v.empty() ? NULL : &v[0], (int64_t)v.size()), "init phase");
    }
    times_[0] += now() - t0;

    for (superstep_ = 0; superstep_ < maxSupersteps_; ++superstep_) {
      int64_t active = 0;
      device_check(gre_engine_step(dev_, &active), "superstep");
      lastCounts_.push_back(active);
      if (binding_.reduceReadsVertexData) download(GRE_ARR_VERTEX, vertexCode_.vertex_data);
      vertexCode_.reduce(dc_);
      if (binding_.usesLevel) device_check(gre_engine_set_level(dev_, levelOf(vertexCode_, 0)), "set level");
      if (dc_.procID() == 0)
        std::cout << "Superstep-" << superstep_ << ": Number of active vertices..." << active << std::endl;
      if (active == 0) break;
    }
    double dt[5];
    device_check(gre_engine_phase_times(dev_, dt), "phase times");
    times_[2] += dt[2];
    times_[4] += dt[4];
    downloadState();
    if (binding_.publishesState) graph_.setProgramState(binding_.id, vertexCode_.vertex_data);
    vertexCode_.output_statistics(dc_);
  }

  // The device program this engine runs (GRE_PROG_*; -1 before the first run
  // of an unbound class).
  int deviceProgram() const { return binding_.id; }

  double getInitVertexTime() { return times_[0]; }
  double getSyncScatterTime() { return times_[1]; }
  double getLocalComputeTime() { return times_[2]; }
  double getSyncCombinerTime() { return times_[3]; }
  double getLocalApplyTime() { return times_[4]; }
  int getSuperstepCount() { return (int)superstep_; }

  // GRE's message entry point, as a program's scatter() calls it
  // (Engine.hpp:1377-1389): combine under the vertex's lock, then mark the
  // destination in nextSet.  The device superstep does not call scatter(), so
  // this only serves host-side use of a program.
  template <typename Ctx_t, typename Data_t>
  void sendMessage(Ctx_t& ctx, lvid_t lv, Data_t& data) {
    typename Ctx_t::Lock_t& lock = ctx.vlock.getLockRef(lv);
    if (vertexCode_.combine(lv, data, lock)) ctx.nextSet.testAndSet(lv);
  }

 private:
  static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  }

  void ensureDevice() {
    if (dev_) return;
    device_check(gre_engine_create(&dev_, graph_.device(), binding_.id), "gre_engine_create");
  }

  // currLevel of programs that keep one (BC): getCurrLevel() when the class has it
  template <class V>
  static auto levelOf(V& v, int) -> decltype((int32_t)v.getCurrLevel()) { return (int32_t)v.getCurrLevel(); }
  template <class V>
  static int32_t levelOf(V&, long) { return 1; }

  // a program that loads its own edge weights (GRE's SSSP: loadEdgeProperty())
  template <class V>
  static auto loadWeights(V& v, int) -> decltype(v.loadEdgeProperty(), void()) { v.loadEdgeProperty(); }
  template <class V>
  static void loadWeights(V&, long) {}

  // ---- the binding check ------------------------------------------------------
  // GRE's superstep at one process (Engine.hpp:246-406, 437-471, 592-625,
  // 757-798 in the GRE tree) with the program's own methods, on the host:
  // master pre-insert of the sources, batches of 8 active vertices (Out edges,
  // then In edges, then assert_to_halt), apply on the vertices that got a
  // message, the active vector rebuilt from the whole bitmap, reduce(dc).
  // Used only on the probe graph.  Returns the per-superstep active counts.
  template <Dir dir>
  std::vector<int64_t> runOnHost(SourceVertexSet_t& srcSet) {
    if (!isInitialized_) initialize();
    const lvid_t n = (lvid_t)graph_.get_num_masters();
    LocalGraph_t& lg = graph_.refLocalGraph();
    VertexSet_t nextSet((int)n);
    Util::vLock_ID locks;
    TP_Context_t ctx(nextSet, locks, this);
    std::vector<char> curr(n, 0);
    std::vector<lvid_t> queue;
    if (srcSet.isALL()) {
      for (lvid_t v = 0; v < n; ++v)
        if (vertexCode_.init(v)) { curr[v] = 1; queue.push_back(v); }
    } else {
      for (size_t i = 0; i < srcSet.vertices.size(); ++i) {
        const lvid_t s = srcSet.vertices[i];
        if (!curr[s]) { curr[s] = 1; queue.push_back(s); }
      }
      for (size_t i = 0; i < srcSet.vertices.size(); ++i) (void)vertexCode_.init(srcSet.vertices[i]);
    }
    std::vector<int64_t> counts;
    for (long ss = 0; ss < maxSupersteps_; ++ss) {
      nextSet.clear_bitmap();
      for (size_t b = 0; b < queue.size(); b += 8) {
        const size_t e = b + 8 < queue.size() ? b + 8 : queue.size();
        if (dir == Out || dir == InOut)
          for (size_t i = b; i < e; ++i) {
            const lvid_t u = queue[i];
            for (typename LocalGraph_t::edge_iterator it = lg.edge_begin(u); it != lg.edge_end(u); ++it)
              vertexCode_.scatter(ctx, u, *it, it);
          }
        if (dir == In || dir == InOut)
          for (size_t i = b; i < e; ++i) {
            const lvid_t u = queue[i];
            for (typename LocalGraph_t::redge_iterator it = lg.redge_begin(u); it != lg.redge_end(u); ++it)
              vertexCode_.scatter(ctx, u, *it, it);
          }
        for (size_t i = b; i < e; ++i)
          if (vertexCode_.assert_to_halt(queue[i])) curr[queue[i]] = 0;
      }
      for (lvid_t v = 0; v < n; ++v)
        if (nextSet.test(v) && vertexCode_.apply(v)) curr[v] = 1;
      queue.clear();
      for (lvid_t v = 0; v < n; ++v)
        if (curr[v]) queue.push_back(v);
      vertexCode_.reduce(dc_);
      counts.push_back((int64_t)queue.size());
      if (queue.empty()) break;
    }
    return counts;
  }

  // First difference between two engines' program arrays and counts, or "".
  static std::string firstDifference(SELF_t& h, SELF_t& d, const std::vector<int64_t>& hc,
                                     const std::vector<int64_t>& dc) {
    std::ostringstream os;
    if (hc != dc) {
      os << "active counts differ (host " << hc.size() << " supersteps, device " << dc.size() << ")";
      for (size_t i = 0; i < hc.size() && i < dc.size(); ++i)
        if (hc[i] != dc[i]) { os << ", first at superstep " << i << ": " << hc[i] << " vs " << dc[i]; break; }
      return os.str();
    }
    const size_t n = h.graph_.get_num_masters();
    VertexCode& a = h.vertexCode_;
    VertexCode& b = d.vertexCode_;
    if (a.vertex_data && b.vertex_data)
      for (size_t v = 0; v < n; ++v)
        if (!Device::sameElem(a.vertex_data[v], b.vertex_data[v])) {
          os << "vertex data of vertex " << v << " differs";
          return os.str();
        }
    if (a.combine_data && b.combine_data)
      for (size_t v = 0; v < n; ++v)
        if (!Device::sameElem(a.combine_data[v], b.combine_data[v])) {
          os << "combine data of vertex " << v << " differs";
          return os.str();
        }
    if (a.scatter_data && b.scatter_data)
      for (size_t v = 0; v < n; ++v)
        if (!Device::sameElem(a.scatter_data[v], b.scatter_data[v])) {
          os << "scatter data of vertex " << v << " differs";
          return os.str();
        }
    return "";
  }

  // Runs the class on the probe graph twice, its own methods on the host and
  // device program b on the GPU, and returns "" when they agree, else what
  // differs.
  template <Dir dir>
  std::string probe(const Device::Binding& b, bool allSources) {
    typedef Device::ProbeMaker<VertexCode, DistributedGraph_t> Maker;
    return probeWith<dir>(b, allSources, std::integral_constant<int, (int)Maker::kind>());
  }
  template <Dir dir>
  std::string probeWith(const Device::Binding&, bool, std::integral_constant<int, Device::ProbeNone>) {
    return "SKIP: the class cannot be constructed from the graph alone";
  }
  template <Dir dir, int K>
  std::string probeWith(const Device::Binding& b, bool allSources, std::integral_constant<int, K>) {
    typedef typename Device::ProbeMaker<VertexCode, DistributedGraph_t>::Holder Holder;
    DistributedGraph_t& pg = Device::probeGraph<DistributedGraph_t>();
    const lvid_t n = (lvid_t)pg.get_num_masters();
    Device::QuietCout quiet;
    std::vector<int32_t> fwd;   // BC level/path-count records of a Forward run from vertex 0
    int32_t nLevels = 0;
    if (K == Device::ProbeBackwardLevels) {
      if (!b.needsForwardState) return "the class is probed as BC Backward, which this device program is not";
      gre_engine* f = NULL;
      device_check(gre_engine_create(&f, pg.device(), GRE_PROG_BC_FORWARD), "probe forward");
      const uint32_t s0 = 0;
      int64_t ss = 0;
      int rc = gre_engine_set_level(f, 1);
      if (rc == 0) rc = gre_engine_run(f, GRE_DIR_OUT, &s0, 1, NULL, 0, &ss);
      fwd.assign(2 * (size_t)n, 0);
      if (rc == 0) rc = gre_engine_download(f, GRE_ARR_VERTEX, &fwd[0]);
      if (rc == 0) rc = gre_engine_get_level(f, &nLevels);
      gre_engine_destroy(f);
      device_check(rc, "probe forward");
      pg.setProgramState(GRE_PROG_BC_FORWARD, &fwd[0]);
    }
    Holder ha(pg, fwd.empty() ? NULL : &fwd[0]);
    Holder hb(pg, fwd.empty() ? NULL : &fwd[0]);
    SELF_t ea(pg, dc_.refUnderlying(), probePool(), ha.vc);
    SELF_t eb(pg, dc_.refUnderlying(), probePool(), hb.vc);
    eb.binding_ = b;
    const long maxSS = maxSupersteps_ < 64 ? maxSupersteps_ : 64;
    ea.configMaxSupersteps(maxSS);
    eb.configMaxSupersteps(maxSS);
    loadWeights(ha.vc, 0);
    loadWeights(hb.vc, 0);
    ea.initialize();
    eb.initialize();
    // identical starting arrays, including entries alloc() left unset
    VertexCode& A = ha.vc;
    VertexCode& B = hb.vc;
    const size_t na = Util::align(n);
    if (A.scatter_data && B.scatter_data) std::memcpy(B.scatter_data, A.scatter_data, sizeof(ScatterData_t) * na);
    if (A.combine_data && B.combine_data && (const void*)B.combine_data != (const void*)B.vertex_data)
      std::memcpy(B.combine_data, A.combine_data, sizeof(CombineData_t) * na);
    if (A.vertex_data && B.vertex_data) std::memcpy(B.vertex_data, A.vertex_data, sizeof(VertexData_t) * n);
    std::vector<int64_t> hc, dc;
    std::string diff;
    try {
      if (K == Device::ProbeBackwardLevels) {
        // GRE's BC driver: Backward run<In> over every vertex, level by level
        SourceVertexSet_t all;
        all.setALL();
        for (int32_t l = nLevels - 2; l >= 0 && diff.empty(); --l) {
          setLevel(ha.vc, l, 0);
          setLevel(hb.vc, l, 0);
          const std::vector<int64_t> h1 = ea.template runOnHost<In>(all);
          eb.template run<In>(all);
          hc.insert(hc.end(), h1.begin(), h1.end());
          dc.insert(dc.end(), eb.lastCounts_.begin(), eb.lastCounts_.end());
        }
      } else {
        SourceVertexSet_t src;
        if (allSources) src.setALL();
        else src.vertices.push_back(0);
        hc = ea.template runOnHost<dir>(src);
        eb.template run<dir>(src);
        dc = eb.lastCounts_;
      }
      diff = firstDifference(ea, eb, hc, dc);
    } catch (const DeviceError& e) {
      diff = std::string("the device run failed: ") + e.what();
    }
    pg.setProgramState(GRE_PROG_BC_FORWARD, NULL);   // a probe Forward may have left its (freed) state
    ea.finalize();
    eb.finalize();
    return diff;
  }
  template <class V>
  static auto setLevel(V& v, int32_t l, int) -> decltype(v.setCurrLevel(l), void()) { v.setCurrLevel(l); }
  template <class V>
  static void setLevel(V&, int32_t, long) {}
  static Util::BufferPool& probePool() {
    static Util::BufferPool pool;
    return pool;
  }

  // The binding this class runs with, checked once per class and direction:
  // its ProgramTraits binding if it has one (which the probe must confirm),
  // else the first device program with matching element sizes that
  // reproduces it on the probe graph.
  template <Dir dir>
  Device::Binding checkedBinding(bool allSources) {
    static std::map<int, Device::Binding> done;
    const int key = (int)dir * 2 + (allSources ? 1 : 0);
    typename std::map<int, Device::Binding>::iterator it = done.find(key);
    if (it != done.end()) return it->second;
    std::ostringstream why;
    Device::Binding pick;
    pick.id = -1;
    if (Traits::id >= 0) {
      const Device::Binding b = Device::bindingFromTraits<Traits>();
      const std::string d = probe<dir>(b, allSources);
      if (d.compare(0, 5, "SKIP:") == 0) {
        std::cerr << "GRE: binding of " << typeid(VertexCode).name() << " to device program " << b.id
                  << " not probed (" << d.substr(6) << ")" << std::endl;
        pick = b;
      } else if (!d.empty()) {
        why << "vertex program " << typeid(VertexCode).name() << " does not compute the rules of device program "
            << b.id << " it is bound to: on the probe graph " << d;
        throw DeviceError(GRE_FAILURE, why.str());
      } else {
        pick = b;
      }
    } else {
      why << "vertex program " << typeid(VertexCode).name() << " has no device binding and matches no device program:";
      for (int id = 0; id < Device::kNumDevicePrograms && pick.id < 0; ++id) {
        Device::Binding b = Device::deviceProgram(id);
        if (b.scatterBytes != (int)sizeof(ScatterData_t) || b.combineBytes != (int)sizeof(CombineData_t) ||
            b.vertexBytes != (int)sizeof(VertexData_t))
          continue;
        b.reduceReadsVertexData = true;  // its reduce() may read the vertex data
        const std::string d = probe<dir>(b, allSources);
        if (d.empty()) pick = b;
        else why << " [" << id << ": " << d << "]";
      }
      if (pick.id < 0) throw DeviceError(GRE_FAILURE, why.str());
    }
    done[key] = pick;
    return pick;
  }

  template <typename T>
  void upload(int which, T* host) {
    if (host) device_check(gre_engine_upload(dev_, which, host), "upload");
  }
  template <typename T>
  void download(int which, T* host) {
    if (host) device_check(gre_engine_download(dev_, which, host), "download");
  }

  void uploadState() {
    upload(GRE_ARR_SCATTER, vertexCode_.scatter_data);
    upload(GRE_ARR_COMBINE, vertexCode_.combine_data);
    upload(GRE_ARR_VERTEX, vertexCode_.vertex_data);
    if (binding_.needsEdgeWeights) {
      const void* w = graph_.edgeProperty();
      if (!w || graph_.edgePropertyBytes() != 4)
        throw DeviceError(GRE_NULL, "this program needs a 4-byte edge property: call loadEdgeProperty first");
      if (w != weightsFrom_) {
        device_check(gre_graph_set_weights_u32(graph_.device(), static_cast<const uint32_t*>(w), 0), "edge weights");
        weightsFrom_ = w;
      }
    }
    if (binding_.needsForwardState) {
      void* fs = graph_.programState(GRE_PROG_BC_FORWARD);
      if (!fs) throw DeviceError(GRE_NULL, "Backward needs a finished Forward run on the same graph");
      device_check(gre_engine_upload(dev_, GRE_ARR_AUX, fs), "upload forward state");
    }
  }
  void downloadState() {
    download(GRE_ARR_SCATTER, vertexCode_.scatter_data);
    if ((const void*)vertexCode_.combine_data != (const void*)vertexCode_.vertex_data)
      download(GRE_ARR_COMBINE, vertexCode_.combine_data);
    download(GRE_ARR_VERTEX, vertexCode_.vertex_data);
  }

  DistributedGraph_t& graph_;
  VertexCode& vertexCode_;
  BufComm_t dc_;
  Device::Binding binding_;
  std::vector<int64_t> lastCounts_;
  gre_engine* dev_;
  bool isInitialized_;
  long maxSupersteps_;
  long superstep_;
  const void* weightsFrom_;
  double times_[5];
};

}  // namespace GRE
#endif
