// Device binding of GRE's vertex programs.
//
// A GRE vertex program is a C++ class whose scatter/combine/apply run inside
// GRE's worker threads.  On the GPU the superstep is a fixed set of
// hand-written gfx950 kernels reached through the C ABI (gre_dev.h), so each
// program is bound to the kernel set that computes exactly its
// init/scatter/combine/apply/assert_to_halt rules (GRE_PROG_*).  The binding
// is a specialisation of ProgramTraits keyed on the program's class; the
// shipped programs (include/Kernels in the GRE tree: PageRank, PageRank_delta,
// BFS, ConnectedComponents, SSSP, Forward, Backward, CountIn) are bound here,
// so they and their drivers compile unchanged.
//
// A binding names rules; it does not prove that a class computes them.  At a
// class's first run<dir>() the engine runs the class's own init / scatter /
// combine / apply / assert_to_halt / reduce on the host over a small probe
// graph, runs the bound device program on the same graph, and throws
// DeviceError when the two differ (Engine/Engine.hpp).  A class with no
// binding is probed against every device program whose element sizes match
// and runs on the first one that reproduces it; if none does, run<dir>()
// throws.  Nothing falls back to a CPU superstep.
//
// Per binding:
//   id                      device program (GRE_PROG_*)
//   scatterBytes/...        element sizes of the three arrays, checked against
//                           the class's typedefs at compile time
//   reduceReadsVertexData   the program's reduce() reads its vertex array on
//                           the host every superstep (ConnectedComponents counts
//                           components), so the engine refreshes it first
//   needsEdgeWeights        scatter adds a per-edge weight (SSSP): the graph's
//                           most recently loaded edge property
//   usesLevel               apply/init compare against currLevel (BC): the
//                           engine pushes level(vc) before every superstep
//   publishesState          the vertex array is left on the graph for a later
//                           program (BC Forward -> Backward)
//   needsForwardState       reads the Forward program's vertex array (Backward)
#ifndef GRE_AMD_DEVICE_PROGRAMS_HPP
#define GRE_AMD_DEVICE_PROGRAMS_HPP
#include <stdint.h>

#include <type_traits>

#include "gre_dev.h"

class PageRank;
class PageRank_delta;
class BFS;
class ConnectedComponents;
class SSSP;
class Forward;
class Backward;
class CountIn;

namespace GRE {
namespace Device {

struct TraitsBase {
  enum {
    reduceReadsVertexData = 0,
    needsEdgeWeights = 0,
    usesLevel = 0,
    publishesState = 0,
    needsForwardState = 0
  };
  template <class VC>
  static int32_t level(VC&) { return 1; }
};

template <class VC>
struct ProgramTraits : TraitsBase {
  enum { id = -1, scatterBytes = 0, combineBytes = 0, vertexBytes = 0 };
};

template <>
struct ProgramTraits< ::PageRank> : TraitsBase {                 // PageRank.hpp:18-126
  enum { id = GRE_PROG_PAGERANK, scatterBytes = 8, combineBytes = 8, vertexBytes = 8 };
};
template <>
struct ProgramTraits< ::PageRank_delta> : TraitsBase {           // PageRank.hpp:129-234
  enum { id = GRE_PROG_PAGERANK_DELTA, scatterBytes = 8, combineBytes = 8, vertexBytes = 8 };
};
template <>
struct ProgramTraits< ::BFS> : TraitsBase {                      // BFS.hpp:18-86
  enum { id = GRE_PROG_BFS, scatterBytes = 8, combineBytes = 8, vertexBytes = 8 };
};
template <>
struct ProgramTraits< ::ConnectedComponents> : TraitsBase {      // CC.hpp:18-125
  enum { id = GRE_PROG_CC, scatterBytes = 8, combineBytes = 8, vertexBytes = 8 };
  enum { reduceReadsVertexData = 1 };
};
template <>
struct ProgramTraits< ::SSSP> : TraitsBase {                     // SSSP.hpp:19-96
  enum { id = GRE_PROG_SSSP, scatterBytes = 4, combineBytes = 4, vertexBytes = 4 };
  enum { needsEdgeWeights = 1 };
};
template <>
struct ProgramTraits< ::Forward> : TraitsBase {                  // BC.hpp:22-95
  enum { id = GRE_PROG_BC_FORWARD, scatterBytes = 4, combineBytes = 4, vertexBytes = 8 };
  enum { usesLevel = 1, publishesState = 1 };
  template <class VC>
  static int32_t level(VC& v) { return v.getCurrLevel(); }
};
template <>
struct ProgramTraits< ::Backward> : TraitsBase {                 // BC.hpp:97-168
  enum { id = GRE_PROG_BC_BACKWARD, scatterBytes = 8, combineBytes = 8, vertexBytes = 8 };
  enum { usesLevel = 1, needsForwardState = 1 };
  template <class VC>
  static int32_t level(VC& v) { return v.getCurrLevel(); }
};
template <>
struct ProgramTraits< ::CountIn> : TraitsBase {                  // CountIn.hpp:18-106
  enum { id = GRE_PROG_COUNTIN, scatterBytes = 4, combineBytes = 4, vertexBytes = 4 };
};

// Runtime form of a binding (also what an unbound class is matched against).
struct Binding {
  int id;
  int scatterBytes, combineBytes, vertexBytes;
  bool reduceReadsVertexData, needsEdgeWeights, usesLevel, publishesState, needsForwardState;
};

template <class T>
inline Binding bindingFromTraits() {
  Binding b;
  b.id = T::id;
  b.scatterBytes = T::scatterBytes;
  b.combineBytes = T::combineBytes;
  b.vertexBytes = T::vertexBytes;
  b.reduceReadsVertexData = T::reduceReadsVertexData != 0;
  b.needsEdgeWeights = T::needsEdgeWeights != 0;
  b.usesLevel = T::usesLevel != 0;
  b.publishesState = T::publishesState != 0;
  b.needsForwardState = T::needsForwardState != 0;
  return b;
}

// The device programs, in GRE_PROG_* order.
inline Binding deviceProgram(int id) {
  switch (id) {
    case GRE_PROG_PAGERANK: return bindingFromTraits<ProgramTraits< ::PageRank> >();
    case GRE_PROG_PAGERANK_DELTA: return bindingFromTraits<ProgramTraits< ::PageRank_delta> >();
    case GRE_PROG_BFS: return bindingFromTraits<ProgramTraits< ::BFS> >();
    case GRE_PROG_CC: return bindingFromTraits<ProgramTraits< ::ConnectedComponents> >();
    case GRE_PROG_SSSP: return bindingFromTraits<ProgramTraits< ::SSSP> >();
    case GRE_PROG_BC_FORWARD: return bindingFromTraits<ProgramTraits< ::Forward> >();
    case GRE_PROG_BC_BACKWARD: return bindingFromTraits<ProgramTraits< ::Backward> >();
    case GRE_PROG_COUNTIN: return bindingFromTraits<ProgramTraits< ::CountIn> >();
    default: {
      Binding b = bindingFromTraits<ProgramTraits<void> >();
      return b;
    }
  }
}
const int kNumDevicePrograms = 8;

// ---- probe instances --------------------------------------------------------
// How the binding check builds two instances of a program class on the probe
// graph, and which runs it compares.  The default takes classes constructible
// from the graph alone and makes one run<dir>() like the caller's.  GRE's
// Forward and Backward take extra constructor arguments; their probes follow
// GRE's BC driver (usr/test_bc.cpp:61-110): Forward from a source, Backward per
// level over the Forward state of the same graph.

// A typed view of an untyped pointer, for constructor arguments whose type
// this header cannot name (GRE's vtxData*).
struct AnyPtr {
  void* p;
  template <class P>
  operator P*() const {
    static_assert(sizeof(P) == 8, "a BC level/path-count record is two 32-bit ints");
    return static_cast<P*>(p);
  }
};

enum ProbeKind { ProbeNone, ProbeOneRun, ProbeBackwardLevels };

template <class VC, class DG, class = void>
struct ProbeMaker {
  enum { kind = ProbeNone };
};
template <class VC, class DG>
struct ProbeMaker<VC, DG, typename std::enable_if<std::is_constructible<VC, DG&>::value>::type> {
  enum { kind = ProbeOneRun };
  struct Holder {
    VC vc;
    Holder(DG& g, void*) : vc(g) {}
  };
};
template <class DG>
struct ProbeMaker< ::Forward, DG, void> {
  enum { kind = ProbeOneRun };
  template <class VC>
  struct HolderT {
    typename VC::VertexData_t* state;
    VC vc;
    HolderT(DG& g, void*) : state(NULL), vc(g, state, 1) {}
    ~HolderT() { delete[] state; }
  };
  typedef HolderT< ::Forward> Holder;
};
template <class DG>
struct ProbeMaker< ::Backward, DG, void> {
  enum { kind = ProbeBackwardLevels };
  template <class VC>
  struct HolderT {
    double* deps;
    VC vc;
    HolderT(DG& g, void* fwdState) : deps(NULL), vc(g, AnyPtr{fwdState}, deps, 0) {}
    ~HolderT() { delete[] deps; }
  };
  typedef HolderT< ::Backward> Holder;
};

}  // namespace Device
}  // namespace GRE
#endif
