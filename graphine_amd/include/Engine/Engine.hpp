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
// and calls output_statistics(dc).  The program's own scatter/combine/apply
// methods are not executed: the device kernels bound to the program in
// Engine/DevicePrograms.hpp compute the same rules.
#ifndef GRE_AMD_ENGINE_HPP
#define GRE_AMD_ENGINE_HPP
#include <stdint.h>

#include <chrono>
#include <iostream>
#include <stdexcept>
#include <string>
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
    static_assert(Traits::id >= 0,
                  "this vertex program has no device binding: add a GRE::Device::ProgramTraits specialisation");
    static_assert(sizeof(ScatterData_t) == (size_t)Traits::scatterBytes, "scatter data type does not match the binding");
    static_assert(sizeof(CombineData_t) == (size_t)Traits::combineBytes, "combine data type does not match the binding");
    static_assert(sizeof(VertexData_t) == (size_t)Traits::vertexBytes, "vertex data type does not match the binding");
    static_assert(dir == Out || dir == In || dir == InOut, "run<dir> needs Out, In or InOut");
    (void)mode;  // both GRE thread modes map to the same device superstep
    if (!isInitialized_) initialize();
    ensureDevice();
    device_check(gre_engine_config_max_supersteps(dev_, maxSupersteps_), "configMaxSupersteps");
    uploadState();
    if (Traits::usesLevel) device_check(gre_engine_set_level(dev_, Traits::level(vertexCode_)), "set level");

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
      if (Traits::reduceReadsVertexData) download(GRE_ARR_VERTEX, vertexCode_.vertex_data);
      vertexCode_.reduce(dc_);
      if (Traits::usesLevel) device_check(gre_engine_set_level(dev_, Traits::level(vertexCode_)), "set level");
      if (dc_.procID() == 0)
        std::cout << "Superstep-" << superstep_ << ": Number of active vertices..." << active << std::endl;
      if (active == 0) break;
    }
    double dt[5];
    device_check(gre_engine_phase_times(dev_, dt), "phase times");
    times_[2] += dt[2];
    times_[4] += dt[4];
    downloadState();
    if (Traits::publishesState) graph_.setProgramState(Traits::id, vertexCode_.vertex_data);
    vertexCode_.output_statistics(dc_);
  }

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
    device_check(gre_engine_create(&dev_, graph_.device(), Traits::id), "gre_engine_create");
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
    if (Traits::needsEdgeWeights) {
      const void* w = graph_.edgeProperty();
      if (!w || graph_.edgePropertyBytes() != 4)
        throw DeviceError(GRE_NULL, "this program needs a 4-byte edge property: call loadEdgeProperty first");
      if (w != weightsFrom_) {
        device_check(gre_graph_set_weights_u32(graph_.device(), static_cast<const uint32_t*>(w), 0), "edge weights");
        weightsFrom_ = w;
      }
    }
    if (Traits::needsForwardState) {
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
  gre_engine* dev_;
  bool isInitialized_;
  long maxSupersteps_;
  long superstep_;
  const void* weightsFrom_;
  double times_[5];
};

}  // namespace GRE
#endif
