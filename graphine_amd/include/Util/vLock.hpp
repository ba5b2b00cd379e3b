// Striped vertex locks (include/Util/vLock.hpp:75-99 in the GRE tree): 8192
// locks indexed by vertex id & 8191.  Kept for source compatibility of vertex
// programs; the device superstep combines with atomics instead.
#ifndef GRE_AMD_VLOCK_HPP
#define GRE_AMD_VLOCK_HPP
#include <cstddef>

#include "Types.hpp"
#include "Util/Atomic.hpp"
#include "Util/Locks.hpp"

#define lockSize 8192L
#define lockMask 8191UL

namespace GRE {
namespace Util {

class vLock_ID {
 public:
  typedef xcasLock lock_t;
  bool tryLock(size_t vid) { return locks_[vid & lockMask].tryLock(); }
  void spinLock(size_t vid) { locks_[vid & lockMask].spinLock(); }
  void unLock(size_t vid) { locks_[vid & lockMask].unLock(); }
  lock_t& getLockRef(size_t vid) { return locks_[vid & lockMask]; }

 private:
  lock_t locks_[lockSize];
};

}  // namespace Util
}  // namespace GRE
#endif
