// Shared host/device helpers of libgre_dev.so (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <vector>
#include <stdexcept>

#include "gre_dev.h"

namespace gre {

// ---- errors ---------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void set_last_error(const std::string& m);

#define GRE_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      throw ::gre::Error(GRE_DEV_FAILURE, std::string(#call) + ": " +              \
                                              hipGetErrorString(e_) + " at " +     \
                                              __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

#define GRE_CHECK(cond, code, msg)                                                 \
  do {                                                                             \
    if (!(cond)) throw ::gre::Error((code), (msg));                                \
  } while (0)

// Runs f, converting exceptions into GRE status codes (no exception crosses the ABI).
template <typename F>
int guard(F&& f) {
  try {
    f();
    return GRE_DEV_SUCCESS;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return GRE_DEV_NOBUFFER;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return GRE_DEV_FAILURE;
  }
}

// ---- runtime state ----------------------------------------------------------
hipStream_t stream();
void ensure_init();

// device memory owned by a handle
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n);
  void release();
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

// scratch reused across launches (grows, never shrinks)
void* scratch(int slot, size_t bytes);

// ---- profiling: HIP events around named launches ----------------------------
struct ProfScope {
  int id;
  hipEvent_t a = nullptr, b = nullptr;
  explicit ProfScope(const char* name);
  ~ProfScope();
};
bool prof_on();

// Appends v to the queue q (length counter qn) for every lane with take set,
// with one atomic per wave (ballot + prefix popcount); order within a wave is
// by lane, across waves by atomic order.
__device__ inline void wave_append(bool take, uint32_t v, uint32_t* __restrict__ q, uint32_t* __restrict__ qn) {
  const unsigned long long bal = __ballot(take);
  if (!bal) return;
  const int lane = (int)(threadIdx.x & 63);
  const int leader = __ffsll(bal) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(qn, (uint32_t)__popcll(bal));
  base = __shfl(base, leader, 64);
  if (take) q[base + (uint32_t)__popcll(bal & ((1ULL << lane) - 1ULL))] = v;
}

inline unsigned grid_for(uint64_t n, unsigned block) {
  uint64_t g = (n + block - 1) / block;
  if (g == 0) g = 1;
  if (g > 0x7fffffffULL) g = 0x7fffffffULL;
  return (unsigned)g;
}

// ---- hashing shared with the R-MAT generator -------------------------------
#define GRE_K0 0x9E3779B97F4A7C15ULL
#define GRE_K1 0xBF58476D1CE4E5B9ULL
#define GRE_K2 0x94D049BB133111EBULL
#define GRE_K3 0x2545F4914F6CDD1DULL
#define GRE_K4 0xD6E8FEB86659FD93ULL

__host__ __device__ inline uint64_t mix64(uint64_t k) {
  k ^= k >> 33; k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

// ---- wave helpers (64 lanes) ------------------------------------------------
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline uint32_t lane_id() { return __lane_id(); }

// atomics used by combine
__device__ inline void atomic_add_f64(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ inline void atomic_add_f32(float* p, float v) { unsafeAtomicAdd(p, v); }

}  // namespace gre
