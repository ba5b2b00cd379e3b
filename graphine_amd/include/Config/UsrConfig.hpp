// Compile-time limits GRE code refers to (include/Config/UsrConfig.hpp:7-18 in
// the GRE tree).  Buffer sizes are unused: no message buffers exist at one GPU.
#ifndef GRE_AMD_USRCONFIG_HPP
#define GRE_AMD_USRCONFIG_HPP
#include <cstddef>
namespace GRE {
const int MAX_NUM_MACHINES = 64;
const int MAX_NUM_THREADS = 32;
const size_t smallBufSize = 4096;
const size_t middleBufSize = 1 << 20;
const size_t largeBufSize = 1 << 23;
const size_t defaultBufPoolSize = 4096;
}  // namespace GRE
#endif
