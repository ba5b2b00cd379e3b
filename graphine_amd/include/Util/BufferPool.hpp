// GRE's message buffer pool (include/Util/BufferPool.hpp:10-221 in the GRE
// tree).  At one GPU no MPI buffer is ever filled, so the pool only records
// its configuration: config()/finalize() are accepted and allocate nothing
// (the GRE drivers ask for 8 GiB of host buffers they would not use here).
#ifndef GRE_AMD_BUFFERPOOL_HPP
#define GRE_AMD_BUFFERPOOL_HPP
#include <cstddef>

#include "Config/UsrConfig.hpp"

namespace GRE {
namespace Util {

class BufferPool {
 public:
  BufferPool() : bufCount_(0), bufSize_(0), initialized_(false) {}
  bool config(size_t bufCount, size_t bufSize = middleBufSize, bool do_initialize = false) {
    bufCount_ = bufCount;
    bufSize_ = bufSize;
    initialized_ = do_initialize;
    return true;
  }
  bool config() { return config(defaultBufPoolSize, middleBufSize, false); }
  void finalize() { initialized_ = false; }
  void destroy() { finalize(); }
  size_t bufSize() const { return bufSize_; }
  size_t bufCount() const { return bufCount_; }
  size_t capacity() const { return bufCount_ * bufSize_; }

 private:
  size_t bufCount_, bufSize_;
  bool initialized_;
};

}  // namespace Util
}  // namespace GRE
#endif
