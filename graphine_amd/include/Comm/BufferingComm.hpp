// The engine's distributed-control object (COMM::BufComm,
// include/Comm/BufferingComm.hpp:14-352 in the GRE tree), as vertex programs
// see it in reduce()/output_statistics(): refUnderlying(), procID(), nProcs().
// At one process there is no buffered message traffic to manage.
#ifndef GRE_AMD_BUFFERING_COMM_HPP
#define GRE_AMD_BUFFERING_COMM_HPP
#include "Comm/CommUnderlying.hpp"

namespace GRE {
namespace COMM {

template <typename BufferPool_t>
class BufComm {
 public:
  typedef Underlying CommUnderlying_t;
  BufComm(CommUnderlying_t& u, BufferPool_t& pool, bool = true) : u_(u), pool_(pool) {}
  CommUnderlying_t& refUnderlying() { return u_; }
  int procID() const { return u_.rank(); }
  int nProcs() const { return u_.size(); }
  void barrier() { u_.barrier(); }

 private:
  CommUnderlying_t& u_;
  BufferPool_t& pool_;
};

}  // namespace COMM
}  // namespace GRE
#endif
