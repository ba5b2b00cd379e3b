// Host atomics used by GRE vertex programs' statistics code, e.g.
// ConnectedComponents::get_num_components (include/Util/Atomic.hpp:99-125 in
// the GRE tree).  <unistd.h> comes first so the sync() helper below does not
// collide with the POSIX function of the same name.
#ifndef GRE_AMD_ATOMIC_HPP
#define GRE_AMD_ATOMIC_HPP
#include <unistd.h>

#define AtomicFetchAdd(u, v) __sync_fetch_and_add((u), (v))
#define AtomicAddFetch(u, v) __sync_add_and_fetch((u), (v))
#define AtomicFetchOr(u, v) __sync_fetch_and_or((u), (v))
#define AtomicOrFetch(u, v) __sync_or_and_fetch((u), (v))
#define AtomicFetchAnd(u, v) __sync_fetch_and_and((u), (v))
#define AtomicAndFetch(u, v) __sync_and_and_fetch((u), (v))
#define AtomicCAS(u, v, w) __sync_val_compare_and_swap((u), (v), (w))
#define int32_cas(u, v, w) __sync_bool_compare_and_swap((u), (v), (w))
#define int64_fetch_add(u, v) __sync_fetch_and_add((u), (v))

namespace GRE {
namespace Util {
inline void full_barrier() { __sync_synchronize(); }
}  // namespace Util
}  // namespace GRE
#endif
