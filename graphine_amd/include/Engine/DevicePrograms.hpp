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
// so they and their drivers compile unchanged.  A program without a binding
// does not compile against SynchronousEngine (static_assert) rather than
// silently running on the CPU.
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

}  // namespace Device
}  // namespace GRE
#endif
