// GRE's global id types and traversal directions (include/Types.hpp:13-37 in
// the GRE tree), restated so GRE programs compile against this build.
#ifndef GRE_AMD_TYPES_HPP
#define GRE_AMD_TYPES_HPP
#include <stdint.h>
#include <cstddef>
#include <limits>
#include <utility>

typedef char byte;
typedef uint32_t lvid_t;   // local vertex id (== vid at one process)
typedef size_t lidx_t;
typedef uint64_t vid_t;    // global vertex id
typedef size_t idx_t;
typedef size_t leid_t;     // local edge id: position in the CSR edge list
typedef std::pair<lvid_t, lvid_t> ledge_t;
typedef std::pair<vid_t, vid_t> edge_t;
typedef int partID_t;
typedef int subgraph_id_t;

template <typename T>
struct WeightedEdge {
  edge_t edge;
  T property;
};

#define NotALvid 0xffffffffL

enum Dir { Out, In, InOut, None };

namespace GRE {
typedef struct notUsed {} notUsed_t;
}
#endif
