// GRE's per-vertex spin lock (xcasLock, include/Util/Locks.hpp:54-96 in the
// GRE tree).  On the device no lock exists: each message's combine is one
// atomic in HBM.  This host class keeps the Lock_t type user combine() code
// names, and is a working lock for host-side callers.
#ifndef GRE_AMD_LOCKS_HPP
#define GRE_AMD_LOCKS_HPP
#include <stdint.h>

namespace GRE {
namespace Util {

class xcasLock {
 public:
  static const int size = 4;
  xcasLock() : lock_(0) {}
  explicit xcasLock(bool locked) : lock_(locked ? 1 : 0) {}
  bool tryLock() { return lock_ == 0 && __sync_bool_compare_and_swap(&lock_, 0, 1); }
  void spinLock() {
    while (!__sync_bool_compare_and_swap(&lock_, 0, 1)) {
      while (lock_ == 1) {
      }
    }
  }
  void unLock() {
    __sync_synchronize();
    lock_ = 0;
  }
  void init() { lock_ = 0; }
  void reset() { lock_ = 0; }

 private:
  volatile int32_t lock_;
};

}  // namespace Util
}  // namespace GRE
#endif
