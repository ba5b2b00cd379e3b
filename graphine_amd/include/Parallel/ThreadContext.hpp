// The context a vertex program's scatter() receives (ThreadContext,
// include/Parallel/ThreadContext.hpp:8-67 in the GRE tree): the engine to call
// sendMessage on, the nextSet frontier and the vertex locks.  The device
// superstep has no per-thread context; this type keeps the programs' scatter
// templates well-formed for host-side use.
#ifndef GRE_AMD_THREAD_CONTEXT_HPP
#define GRE_AMD_THREAD_CONTEXT_HPP
#include "Util/vLock.hpp"

namespace GRE {

template <typename VertexSet_t, typename Engine_t>
class HostContext {
 public:
  typedef Util::vLock_ID vLock_t;
  typedef vLock_t::lock_t Lock_t;
  HostContext(VertexSet_t& next, vLock_t& locks, Engine_t* e)
      : threadID(0), nGroups(1), groupID(0), nextSet(next), vlock(locks), engine(e) {}
  int threadID;
  int nGroups;
  int groupID;
  VertexSet_t& nextSet;
  vLock_t& vlock;
  Engine_t* engine;
};

}  // namespace GRE
#endif
