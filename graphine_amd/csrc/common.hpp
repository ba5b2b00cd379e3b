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
  void* p = nullptr;     // the buffer (aligned as asked)
  size_t bytes = 0;
  void* raw = nullptr;   // what hipMalloc returned
  size_t held = 0;       // bytes allocated
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), raw(o.raw), held(o.held) {
    o.p = o.raw = nullptr;
    o.bytes = o.held = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p; bytes = o.bytes; raw = o.raw; held = o.held;
      o.p = o.raw = nullptr;
      o.bytes = o.held = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  // align (a power of two; ignored up to 256, hipMalloc's own alignment): the
  // buffer starts on that boundary, at the cost of up to align more bytes
  void alloc(size_t n, size_t align = 0);
  void release();
  // a view of o's memory: never freed through this buffer, valid while o is
  void view(const DevBuf& o) {
    release();
    p = o.p;
    bytes = o.bytes;
  }
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

// Vertex-label scramble of the R-MAT generator (Graph500's generator permutes
// its labels the same way): a bijection of [0, 2^scale) keyed by the seed,
// each step invertible modulo 2^scale (add, odd multiply, xor with a right
// shift of itself).
__host__ __device__ inline uint32_t rmat_scramble(int scale, uint64_t seed, uint32_t x) {
  const uint64_t mask = (1ULL << scale) - 1;
  const uint64_t k = mix64(seed ^ 0xA0761D6478BD642FULL);
  uint64_t y = x;
  y = (y + k) & mask;
  y = (y * ((k >> 32) | 1ULL)) & mask;
  y ^= y >> ((scale + 1) / 2);
  y = (y * GRE_K0) & mask;
  y ^= y >> ((scale + 2) / 3);
  y = (y + (k >> 11)) & mask;
  return (uint32_t)y;
}

// ---- wave helpers (64 lanes) ------------------------------------------------
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline uint32_t lane_id() { return __lane_id(); }
__device__ inline unsigned long long wave_or(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}

// Block-level counter updates.  A counter that every wave of a large grid
// updates serialises in L2 (about 12 ns per update at one address on MI355X,
// so 262,144 waves spend 3 ms on it); these make one update per block.  Every
// thread of the block must call them, the same number of times, with
// blockDim.x a multiple of 64 and at most 1024.

// *p += sum of v over the block (skipped when zero)
template <typename T>
__device__ inline void block_add(T v, T* p) {
  __shared__ T part[16];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    T t = 0;
    for (unsigned w = 0; w < (blockDim.x >> 6); ++w) t += part[w];
    if (t) atomicAdd(p, t);
  }
  __syncthreads();
}

// *pa += sum of a, *pb += sum of b over the block (each skipped when zero)
template <typename A, typename B>
__device__ inline void block_add2(A a, A* pa, B b, B* pb) {
  __shared__ A sa[16];
  __shared__ B sb[16];
  a = wave_sum(a);
  b = wave_sum(b);
  if ((threadIdx.x & 63) == 0) { sa[threadIdx.x >> 6] = a; sb[threadIdx.x >> 6] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    A ta = 0;
    B tb = 0;
    for (unsigned w = 0; w < (blockDim.x >> 6); ++w) { ta += sa[w]; tb += sb[w]; }
    if (ta) atomicAdd(pa, ta);
    if (tb) atomicAdd(pb, tb);
  }
  __syncthreads();
}

// reserves n slots of *qn for the calling thread, one counter update per
// block; returns the first of them
__device__ inline uint32_t block_reserve(uint32_t n, uint32_t* __restrict__ qn) {
  __shared__ uint32_t wtot[16];
  __shared__ uint32_t bbase;
  const int lane = (int)(threadIdx.x & 63);
  const int w = (int)(threadIdx.x >> 6);
  uint32_t x = n;  // inclusive scan over the wave
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wtot[w] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (unsigned i = 0; i < (blockDim.x >> 6); ++i) {
      const uint32_t c = wtot[i];
      wtot[i] = t;
      t += c;
    }
    bbase = t ? atomicAdd(qn, t) : 0u;
  }
  __syncthreads();
  const uint32_t r = bbase + wtot[w] + x - n;
  __syncthreads();
  return r;
}

// appends v to q for every thread with take, one counter update per block
// (order within q is unspecified)
__device__ inline void block_append(bool take, uint32_t v, uint32_t* __restrict__ q, uint32_t* __restrict__ qn) {
  __shared__ uint32_t wcnt[16];
  __shared__ uint32_t bbase;
  const unsigned long long bal = __ballot(take);
  const int lane = (int)(threadIdx.x & 63);
  const int w = (int)(threadIdx.x >> 6);
  if (lane == 0) wcnt[w] = (uint32_t)__popcll(bal);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (unsigned i = 0; i < (blockDim.x >> 6); ++i) {
      const uint32_t c = wcnt[i];
      wcnt[i] = t;
      t += c;
    }
    bbase = t ? atomicAdd(qn, t) : 0u;
  }
  __syncthreads();
  if (take) q[bbase + wcnt[w] + (uint32_t)__popcll(bal & ((1ULL << lane) - 1ULL))] = v;
  __syncthreads();
}


// atomics used by combine
__device__ inline void atomic_add_f64(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ inline void atomic_add_f32(float* p, float v) { unsafeAtomicAdd(p, v); }

}  // namespace gre
