// GRE::Msg (include/Message.hpp:5-14 in the GRE tree).  At one process no
// message crosses a process boundary; the type stays for source compatibility.
#ifndef GRE_AMD_MESSAGE_HPP
#define GRE_AMD_MESSAGE_HPP
#include "Types.hpp"
namespace GRE {
template <typename data_t, typename dest_t = vid_t>
struct Msg {
  dest_t dest;
  data_t data;
  Msg(dest_t d, data_t v) : dest(d), data(v) {}
  Msg() {}
};
}  // namespace GRE
#endif
