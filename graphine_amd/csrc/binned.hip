// Binned forward pull (propagation blocking) of the single-vector FP32
// PageRank iteration, x' = d (W^T x + (g.x) p) + (1 - d) p with W_uv =
// 1/outdeg(u): the same superstep as the tile pulls, with the random accesses
// moved into LDS.
//
// The gathers of a pull cost one cache request per adjacency entry, which is
// what bounds the tile and partitioned pulls (scattered 4-byte reads run at the
// caches' request rate, not at HBM bandwidth).  Here every HBM access is a
// stream:
//   scatter:    chunk w (consecutive source rows, at most kChunkRows of them
//               and kChunkEdges entries) stages its sources' rank/out-degree in
//               LDS, then walks its entries in (bin, destination) order,
//               reading each value from LDS; the entries to one destination
//               (up to 16 consecutive ones) are summed and written as one slot
//               of the bin array, where bin b holds the runs (w, b) of every
//               chunk in chunk order.  A run is ~140 consecutive slots, so the
//               writes stream too.
//   accumulate: each bin (at most kBinRows consecutive destination rows) is
//               summed in LDS: 64-bit fixed-point atomics (exact, so the sum
//               does not depend on the order of the additions), then each
//               row's update.  A bin with more than kSliceEdges slots is cut
//               into slices whose sums meet in 64-bit integers in HBM,
//               finished by the last pass with the rows without in-edges.
// Rows are relabelled so that bins have equal entry counts: the vertices with
// in-edges ranked by in-degree are dealt to the bins in snake order (round k
// left to right when k is even, right to left otherwise), a bin's rows being
// consecutive; the vertices without in-edges follow, those with out-edges first
// by descending out-degree.  Ranks, personalisation and out-degrees live in
// that order during a run (the handle turns the ranks back on first read), and
// contributions are indexed by row, so every row writes its rank and its
// contribution in sequence.
#include "binned.hpp"
#include "pull.hpp"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <vector>

namespace gre {
namespace {

constexpr unsigned kB = 256;
constexpr int kSrcBits = 14;  // source row within a chunk: kChunkRows = 2^14
constexpr int kDstBits = 14;  // destination row within a bin: kBinRows = 2^14
static_assert(kChunkRows == (1 << kSrcBits), "chunk rows are 2^kSrcBits");
static_assert(kBinRows == (1 << kDstBits), "bin rows are 2^kDstBits");
static_assert(kBinMaxBins <= 65536 && kChunkRows <= 65536, "a scatter word holds 16-bit bin and row fields");
static_assert(kBinRows <= 65536, "dloc is 16 bits");
constexpr int kScatterBlock = 1024;
constexpr int kScatterU = 8;     // entries in flight per thread
// the engine's FP64 scatter: 4 entries in flight per thread measured faster
// than 8 (3.50 against 3.79-3.88 ms at R-MAT scale 26), and chunks of 16,384
// rows (a 128 KB stage, one workgroup per CU) faster than 8,192 (two per CU,
// runs half as long: 3.80-4.52 ms)
constexpr int kEngScatterU = 4;
constexpr int kEngChunkRows = 16384;
constexpr int kAccumBlock = 1024;
constexpr int kAccumU = 8;
constexpr uint32_t kTailRows = 16384;  // rows without in-edges per accumulate workgroup after the slices
constexpr uint32_t kBinKey = 1u << 14;  // cache key of the layout among a graph's partitions

unsigned gcap(uint64_t n) {
  uint64_t g = (n + kB - 1) / kB;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

uint64_t align64(uint64_t n) { return (n + 63) & ~63ULL; }

unsigned bits_of(uint64_t n) {
  unsigned b = 1;
  while (b < 63 && (1ULL << b) < n) ++b;
  return b;
}

__device__ inline uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}

// fixed-order sum of two doubles over a block of up to 1024 threads; thread 0
// gets the totals
__device__ inline void block_sum2_big(double& a, double& b, double* red /*[32]*/) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) { red[w] = a; red[16 + w] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = 0.0;
    b = 0.0;
    for (unsigned i = 0; i < (blockDim.x >> 6); ++i) { a += red[i]; b += red[16 + i]; }
  }
}

// ---- iteration kernels ------------------------------------------------------

__device__ inline float fix_to_float(unsigned long long v) {
  return (float)((double)(long long)v * 0x1p-58);
}
static_assert(kBinFix == 58, "fix_to_float and the scatter scale use 2^58");

// The accumulate's fixed-point conversions.  The forward: the fixed scale
// 2^58 (its sums are bounded by the ranks' L1 norm).
struct Fix58 {
  __device__ long long to(float x) const { return __float2ll_rn(x * 0x1p58f); }
  __device__ float from(unsigned long long v) const { return fix_to_float(v); }
};
// A scale 2^T for one launch: with every row's sum of |values| below 2^e,
// T = 61 - e keeps each value, each partial sum and each row's sum below 2^62
// in magnitude, whatever the order of the additions.
struct FixDyn {
  double sc, isc;  // 2^T, 2^-T
  __device__ long long to(float x) const { return __double2ll_rn((double)x * sc); }
  __device__ float from(unsigned long long v) const { return (float)((double)(long long)v * isc); }
};

// Row r's update from its in-edge sum s (the same formula as the tile pulls'
// forward mode, contribution indexed by row: rank / out-degree, or the rank
// itself when the scatter multiplies by normalised edge weights, CRAW).
template <bool R, bool CRAW>
struct BinUpdate {
  float* contrib_out;
  float* x;
  const float* p_r;
  const int32_t* outdeg_r;
  const double* scal_in;
  float damp;
  float omd;   // 1 - damp
  float invn;
  struct Pre {
    float pv, old;
    int32_t d;
  };
  __device__ Pre pre(uint32_t r) const {
    Pre q;
    q.pv = p_r ? p_r[r] : invn;
    q.old = R ? x[r] : 0.f;
    q.d = outdeg_r[r];
    return q;
  }
  __device__ void fin(uint32_t r, float s, const Pre& q, double& a0, double& a1) const {
    const float dang = (float)scal_in[0];
    const float y = damp * (s + dang * q.pv) + omd * q.pv;
    x[r] = y;
    if (q.d > 0) contrib_out[r] = CRAW ? y : y / (float)q.d;
    else a0 += (double)y;
    if (R) a1 += fabs((double)y - (double)q.old);
  }
  __device__ void apply(uint32_t r, float s, double& a0, double& a1) const { fin(r, s, pre(r), a0, a1); }
  __device__ Fix58 fix() const { return {}; }
};

// Row u's adjoint update over the binned layout of the reversed graph (its
// rows are u's out-edges' heads: the same formula as the tile pulls' adjoint
// mode).  s = the sum of lambda_v over u's out-edges (u -> v), times the
// normalised weight w_uv (W); unweighted sums are divided by u's out-degree d:
//   (W lambda)_u = s,  lambda'_u = c_u + damp (s + [d == 0] p.lambda).
// The fixed-point scale follows max |lambda| (the bits in *mx_in) times F, a
// bound on sum_v |w_uv| over every row (the largest out-degree, times the
// largest |w_uv| when weighted).
template <bool W>
struct BinAdjUpdate {
  float* lout;
  float* wl;
  const float* lin;
  const float* c_r;
  const float* p_r;
  const int32_t* deg_r;     // out-degree of each row's vertex
  const double* scal_in;    // [0] = p.lambda of lin
  const uint32_t* mx_in;    // bits of max |lin|
  double F;
  float damp;
  float invn;
  struct Pre {
    float cu, old, pv;
    int32_t d;
  };
  __device__ FixDyn fix() const {
    const double S = (double)__uint_as_float(*mx_in) * F;
    int e = 0;
    if (S > 0.0 && S < 0x1p1000) frexp(S, &e);  // S < 2^e
    return FixDyn{ldexp(1.0, 61 - e), ldexp(1.0, e - 61)};
  }
  __device__ Pre pre(uint32_t r) const {
    Pre q;
    q.cu = c_r[r];
    q.old = lin[r];
    q.pv = p_r ? p_r[r] : invn;
    q.d = deg_r[r];
    return q;
  }
  __device__ void fin(uint32_t r, float s, const Pre& q, double& a0, double& a1) const {
    if (!W && q.d > 0) s = s / (float)q.d;
    wl[r] = s;
    const float y = q.cu + damp * (s + (q.d == 0 ? (float)scal_in[0] : 0.f));
    lout[r] = y;
    a0 += (double)q.pv * (double)y;
    a1 += fabs((double)y - (double)q.old);
  }
  __device__ void apply(uint32_t r, float s, double& a0, double& a1) const { fin(r, s, pre(r), a0, a1); }
};


// Lane l of a wave holds entry 64k + l of its chunk.  Entries of one run that
// go to the same destination are consecutive (a run is sorted by destination)
// and form groups, cut at every 16-lane row of the wave: a group's values are
// summed in FP32 in a fixed order (an inclusive scan over DPP row shifts, at
// most four additions per lane) and its last lane writes the sum to the
// group's slot.  A word holds the entry's source row within the chunk (14
// bits), bit 14 set if the entry starts a run and bit 15 set if it starts a
// group (entry 0 of every row does); per 64-entry block, rcnt holds (runs of
// the chunk before the block + bit 0 of its run starts - 1) | (groups of the
// chunk before the block) << 12.
constexpr int kGroupLanes = 16;   // a DPP row
constexpr uint32_t kTrashBlocks = 256;  // FP64 scatter: workgroups (mod this) with trash slots of their own
constexpr int kRunBits = 12;      // runs of a chunk <= kBinMaxBins
constexpr uint32_t kWordRun = 1u << 14, kWordGroup = 1u << 15;
static_assert(kSrcBits == 14, "a scatter word holds a 14-bit source and two flags");
static_assert(kBinMaxBins <= (1 << kRunBits), "run counts are 12 bits");
static_assert(kChunkEdges <= (1u << (32 - kRunBits)), "group counts are 20 bits");

// lanes whose bit is set in m (the D lanes below are in their group) add the
// value D lanes down within their 16-lane row
template <int D>
__device__ inline float group_step(float s, uint64_t m) {
  const float y =
      __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, s), 0x110 | D, 0xf, 0xf, true));
  return __builtin_amdgcn_inverse_ballot_w64(m) ? s + y : s;
}

// Inclusive scan of every group; c: bit l set if lane l continues lane l - 1's
// group (wave-uniform).  After the step of offset D, lane l holds the sum of
// lanes max(head, l - 2D + 1) .. l of its group; groups lie in one row, so
// four steps finish them.
__device__ inline float group_scan(float s, uint64_t c) {
  s = group_step<1>(s, c);
  c &= c << 1;
  s = group_step<2>(s, c);
  c &= c << 2;
  s = group_step<4>(s, c);
  c &= c << 4;
  return group_step<8>(s, c);
}

// the same for FP64 values (the engine's sums): two 32-bit DPP moves per shift
template <int D>
__device__ inline double row_shr64(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, 0x110 | D, 0xf, 0xf, true);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), 0x110 | D, 0xf, 0xf, true);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
template <int D>
__device__ inline double group_step(double s, uint64_t m) {
  const double y = row_shr64<D>(s);
  return __builtin_amdgcn_inverse_ballot_w64(m) ? s + y : s;
}
__device__ inline double group_scan(double s, uint64_t c) {
  s = group_step<1>(s, c);
  c &= c << 1;
  s = group_step<2>(s, c);
  c &= c << 2;
  s = group_step<4>(s, c);
  c &= c << 4;
  return group_step<8>(s, c);
}

__device__ inline uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

extern "C" __device__ void gre_buffer_store_idx(float v, __amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset,
                                                int soffset, int aux) __asm("llvm.amdgcn.struct.ptr.buffer.store.f32");

// Scatter: workgroup b runs chunk xs[b % 8] + b / 8 (workgroups go to the XCDs
// in turn), so each XCD walks its own contiguous stretch of chunks, and
// consecutive chunks (whose runs of one bin are adjacent in the bin array)
// share one L2.  Entry i of the chunk reads
// its source's value from the LDS stage; the last entry of each group writes
// the group's sum to slot c + rb[rank], c = the chunk's groups before it and
// rank = its run, counted from the wave's flags (ballots) and its block's
// rcnt (a uniform load).
//
// One counter tracks a wave's loads and stores together, so a round that
// waited for its loads behind the previous round's stores would wait for those
// stores too.  The rounds therefore alternate between two register sets (the
// next round's loads are issued before this round's stores), and no store sits
// in a branch: every lane stores, through a buffer resource with one record per
// bin-array slot, and lanes with nothing to write give an index past the last
// record, which the hardware drops.  A resource's byte offsets are 32-bit, so
// a bin array of 2^30 slots or more (BIG) is stored through four resources of
// 2^30 slots each, one store per resource.
template <int U, bool W, bool BIG, typename T = float, int CR = kChunkRows>
__global__ __launch_bounds__(kScatterBlock) void k_bin_scatter(
    const T* __restrict__ contrib, const uint32_t* __restrict__ ch_r0, const uint32_t* __restrict__ ch_nr,
    const uint32_t* __restrict__ ch_e0, const uint32_t* __restrict__ ch_w0, const uint32_t* __restrict__ ch_q0,
    const uint16_t* __restrict__ src16, const uint32_t* __restrict__ rcnt, const uint32_t* __restrict__ rbase,
    const float* __restrict__ wbin, XcdStretches xs, uint32_t nslots, T* __restrict__ binval,
    T* __restrict__ trash) {
  constexpr bool F64 = std::is_same<T, double>::value;
  __shared__ T sv[CR];
  __shared__ uint32_t rb[kBinMaxBins];
  const uint32_t x = blockIdx.x & 7u;
  const uint32_t w = xs.s[x] + (blockIdx.x >> 3);
  if (w >= xs.s[x + 1]) return;
  const uint32_t r0 = ch_r0[w], nr = ch_nr[w];
  const uint32_t n = ch_e0[w + 1] - ch_e0[w];
  const uint32_t w0 = ch_w0[w], q0 = ch_q0[w], nq = ch_q0[w + 1] - q0;
  const uint16_t* __restrict__ sp = src16 + w0;
  const float* __restrict__ wq = W ? wbin + w0 : nullptr;
  const uint32_t* __restrict__ cp = rcnt + (w0 >> 6);
  // FP32: stores through buffer resources; BIG: the bin array is split in four
  // windows of 2^30 slots, one resource each.  FP64 (whose slots pass 4 GiB
  // of offsets sooner): plain stores, and lanes with nothing to write store to
  // this wave's own trash slots (2 U x 64 values), so no store is
  // conditional either
  constexpr int NW = BIG ? 4 : 1;
  __amdgpu_buffer_rsrc_t out[NW];
  if constexpr (!F64) {
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      const uint64_t lo = (uint64_t)k << 30;
      const uint32_t nrec = nslots > lo ? (uint32_t)std::min<uint64_t>(nslots - lo, 1ULL << 30) : 0u;
      out[k] = __builtin_amdgcn_make_buffer_rsrc(binval + lo, (short)4, (int)nrec, 0x00020000);
    }
  }
  T* __restrict__ tw =
      F64 ? trash + ((size_t)(blockIdx.x & (kTrashBlocks - 1)) * (kScatterBlock / 64) + (threadIdx.x >> 6)) * (2 * U * 64)
          : nullptr;
  const uint32_t lane = threadIdx.x & 63u;
  constexpr uint32_t step = kScatterBlock * U;
  struct Set {
    uint16_t wd[U];  // widened only where used, so that no use sits next to the load
    uint32_t c[U];
    float wt[W ? U : 1];
  };
  // lanes past the end take the last entry, which lies in the 64-entry block of
  // the wave's first lane, so the block index is uniform across the wave
  auto load = [&](Set& t, uint32_t base) {
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const uint32_t i = base + j * kScatterBlock + threadIdx.x;
      const uint32_t ii = i < n ? i : n - 1;
      t.wd[j] = __builtin_nontemporal_load(sp + ii);
      if constexpr (W) t.wt[j] = __builtin_nontemporal_load(wq + ii);
      t.c[j] = cp[__builtin_amdgcn_readfirstlane(ii >> 6)];
    }
  };
  // lanes past the end start groups of their own and no run, so the last
  // entry ends a group and every rank stays below the chunk's run count
  auto put = [&](const Set& t, uint32_t base) {
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const uint32_t i = base + j * kScatterBlock + threadIdx.x;
      const uint32_t wd = t.wd[j];
      const uint64_t valid = __builtin_amdgcn_ballot_w64(i < n);
      const uint64_t gs = __builtin_amdgcn_ballot_w64(wd & kWordGroup) | ~valid;
      const uint64_t rs = __builtin_amdgcn_ballot_w64(wd & kWordRun) & valid;
      T x = sv[wd & (kWordRun - 1)];
      if constexpr (W) x *= t.wt[j];
      const T v = group_scan(x, ~gs);
      const uint32_t rank = (t.c[j] & ((1u << kRunBits) - 1)) + lanes_below(rs >> 1);
      const uint32_t at = (t.c[j] >> kRunBits) + lanes_below(gs >> 1) + rb[rank];
      const bool tail = __builtin_amdgcn_inverse_ballot_w64(((gs >> 1) | (1ULL << 63)) & valid);
      if constexpr (F64) {
        *(tail ? binval + at : tw + lane) = v;
      } else if constexpr (BIG) {
#pragma unroll
        for (int k = 0; k < NW; ++k)
          gre_buffer_store_idx(v, out[k], tail && (at >> 30) == (uint32_t)k ? (int)(at & ((1u << 30) - 1)) : -1, 0, 0,
                               0);
      } else {
        gre_buffer_store_idx(v, out[0], tail ? (int)at : -1, 0, 0, 0);
      }
      // one entry at a time: computed all at once, the eight entries' ballots
      // overflow the scalar registers
      if (j + 1 < U) __builtin_amdgcn_sched_barrier(0);
    }
  };
  {
    // the stage's loads are all issued before any is stored (lanes past the
    // end reload the last value); the stores fill both arrays whole, with no
    // branch
    constexpr int FS = CR / kScatterBlock, FR = kBinMaxBins / kScatterBlock;
    T f[FS];
    uint32_t q[FR];
#pragma unroll
    for (int k = 0; k < FS; ++k) {
      const uint32_t i = threadIdx.x + k * kScatterBlock;
      f[k] = contrib[r0 + (i < nr ? i : nr - 1)];
    }
#pragma unroll
    for (int k = 0; k < FR; ++k) {
      const uint32_t i = threadIdx.x + k * kScatterBlock;
      q[k] = rbase[q0 + (i < nq ? i : nq - 1)];
    }
#pragma unroll
    for (int k = 0; k < FS; ++k) sv[threadIdx.x + k * kScatterBlock] = f[k];
#pragma unroll
    for (int k = 0; k < FR; ++k) rb[threadIdx.x + k * kScatterBlock] = q[k];
  }
  __syncthreads();
  // Rounds go in pairs with no exit between a load and its use (a branch there
  // lets the compiler sink the loads below the stores), and the way into the
  // loop ends like the loop's own latch: one set's loads, then two rounds'
  // stores (here dropped ones), so that the waits the compiler places cover
  // no store.  sched_barrier keeps each round's loads above the previous
  // round's stores.
  const uint32_t rounds = (n + step - 1) / step;
  Set a, b;
  load(a, 0);
#pragma unroll
  for (int j = 0; j < 2 * U; ++j) {  // distinct, or they merge
    if constexpr (F64) tw[64 * j + lane] = 0.0;
    else gre_buffer_store_idx(0.f, out[0], -1 - j, 0, 0, 0);
  }
  for (uint32_t r = 0; r + 1 < rounds; r += 2) {
    load(b, (r + 1) * step);
    __builtin_amdgcn_sched_barrier(0);
    put(a, r * step);
    load(a, (r + 2) * step);
    __builtin_amdgcn_sched_barrier(0);
    put(b, (r + 1) * step);
  }
  if (rounds & 1) put(a, (rounds - 1) * step);
}

// Accumulate slice s of bin b: 64-bit fixed-point sums in LDS, then either the
// rows' updates (the bin's only slice) or the sums added into the bin's slots
// of gacc (finished by k_bin_fin).  (acc0, acc1) partials per slice.  Rounds
// alternate between two register sets as in the scatter; lanes past the end
// add zero to the first entry's row.
template <class Up, int U>
__global__ __launch_bounds__(kAccumBlock) void k_bin_accum(
    const float* __restrict__ binval, const uint16_t* __restrict__ dloc, const uint32_t* __restrict__ sl_e0,
    const uint32_t* __restrict__ sl_bin, const uint32_t* __restrict__ sl_g, const uint32_t* __restrict__ bin_r0,
    uint32_t nsl, uint32_t nd, uint32_t nV, Up up, unsigned long long* __restrict__ gacc,
    double* __restrict__ part) {
  __shared__ unsigned long long acc[kBinRows];
  __shared__ double red[32];
  const uint32_t s = blockIdx.x;
  if (s >= nsl) {
    // the workgroups after the slices update kTailRows rows without in-edges
    // each (sum 0; they need nothing from the slices, and run in the CUs the
    // last slices leave idle)
    constexpr int Q = kTailRows / kAccumBlock;
    const uint32_t t0 = nd + (s - nsl) * kTailRows;
    typename Up::Pre q[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      const uint32_t r = t0 + threadIdx.x + k * kAccumBlock;
      q[k] = up.pre(r < nV ? r : nd);
    }
    double a0 = 0.0, a1 = 0.0;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      const uint32_t r = t0 + threadIdx.x + k * kAccumBlock;
      if (r < nV) up.fin(r, 0.f, q[k], a0, a1);
    }
    block_sum2_big(a0, a1, red);
    if (threadIdx.x == 0) { part[2 * s] = a0; part[2 * s + 1] = a1; }
    return;
  }
  const auto fx = up.fix();
  const uint32_t b = sl_bin[s];
  const uint32_t r0 = bin_r0[b], nr = bin_r0[b + 1] - r0;
  const uint32_t e0 = sl_e0[s], n = sl_e0[s + 1] - e0;
  const float* __restrict__ vp = binval + e0;
  const uint16_t* __restrict__ dp = dloc + e0;
  constexpr uint32_t step = kAccumBlock * U;
  auto load = [&](float (&v)[U], uint32_t (&d)[U], uint32_t base) {
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const uint32_t i = base + j * kAccumBlock + threadIdx.x;
      const uint32_t ii = i < n ? i : 0u;
      v[j] = __builtin_nontemporal_load(vp + ii);
      d[j] = __builtin_nontemporal_load(dp + ii);
    }
  };
  auto add = [&](const float (&v)[U], const uint32_t (&d)[U], uint32_t base) {
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const uint32_t i = base + j * kAccumBlock + threadIdx.x;
      const float x = i < n ? v[j] : 0.f;
      atomicAdd(&acc[d[j]], (unsigned long long)fx.to(x));
    }
  };
  float va[U], vb[U];
  uint32_t da[U], db[U];
  load(va, da, 0);  // in flight while the accumulators are cleared
  for (uint32_t i = threadIdx.x; i < nr; i += kAccumBlock) acc[i] = 0ULL;
  __syncthreads();
  for (uint32_t base = 0;;) {
    load(vb, db, base + step);
    add(va, da, base);
    base += step;
    if (base >= n) break;
    load(va, da, base + step);
    add(vb, db, base);
    base += step;
    if (base >= n) break;
  }
  __syncthreads();
  double a0 = 0.0, a1 = 0.0;
  const uint32_t g = sl_g[s];
  if (g == 0xffffffffu) {
    // every row's data is loaded before any row is updated
    constexpr int Q = kBinRows / kAccumBlock;
    typename Up::Pre q[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      const uint32_t i = threadIdx.x + k * kAccumBlock;
      q[k] = up.pre(r0 + (i < nr ? i : 0u));
    }
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      const uint32_t i = threadIdx.x + k * kAccumBlock;
      if (i < nr) up.fin(r0 + i, fx.from(acc[i]), q[k], a0, a1);
    }
  } else {
    for (uint32_t i = threadIdx.x; i < nr; i += kAccumBlock)
      if (acc[i]) atomicAdd(&gacc[g + i], acc[i]);
  }
  block_sum2_big(a0, a1, red);
  if (threadIdx.x == 0) { part[2 * s] = a0; part[2 * s + 1] = a1; }
}

// Finish: the rows of the bins cut into slices (their sums in gacc, which is
// cleared for the next iteration).
template <class Up>
__global__ __launch_bounds__(kB) void k_bin_fin(const uint32_t* __restrict__ mb_r0,
                                                const uint32_t* __restrict__ mb_g, uint32_t nmb, uint32_t nmrows,
                                                Up up, unsigned long long* __restrict__ gacc,
                                                double* __restrict__ part) {
  __shared__ double red[32];
  const auto fx = up.fix();
  double a0 = 0.0, a1 = 0.0;
  for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < nmrows; q += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = nmb;  // largest k with mb_g[k] <= q
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (mb_g[mid] <= q) lo = mid;
      else hi = mid;
    }
    const unsigned long long v = gacc[q];
    gacc[q] = 0ULL;
    up.apply(mb_r0[lo] + (uint32_t)(q - mb_g[lo]), fx.from(v), a0, a1);
  }
  block_sum2_big(a0, a1, red);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = a0; part[2 * blockIdx.x + 1] = a1; }
}

// ---- the engine's FP64 sum pull ----------------------------------------------
//
// Each value v of a slot is added into its row as two 64-bit integers: hi =
// round(v 2^K) and lo = round((v 2^K - hi) 2^L) (both scalings exact, and the
// difference exact), so a row's sums are exact integers, independent of the
// order of the additions; the row's sum is hi 2^-K + lo 2^-(K+L), within
// n 2^-(K+L+1) of the exact sum of its n values.  K and L come from the
// largest |value| m and the largest in-degree n_max: with 2^e >= m n_max and
// 2^f > n_max, K = 61 - e and L = 61 - f keep |sum hi| < 2^62 and
// |sum lo| < 2^61.
__device__ inline void eng_scales(const unsigned long long* __restrict__ smax_bits, uint32_t max_in, int& K, int& L) {
  const double m = __longlong_as_double((long long)*smax_bits);
  int f = 0;
  frexp((double)max_in, &f);
  L = 61 - f;
  if (!(m > 0.0)) {
    K = 0;
    return;
  }
  int e = 0;
  frexp(m * (double)max_in, &e);
  K = 61 - e;
}

__device__ inline void eng_split(double x, int K, int L, long long& hi, long long& lo) {
  const double xs = ldexp(x, K);
  const double h = rint(xs);
  hi = (long long)h;
  lo = (long long)rint(ldexp(xs - h, L));
}

__device__ inline double eng_join(unsigned long long hi, unsigned long long lo, int K, int L) {
  return ldexp((double)(long long)hi, -K) + ldexp((double)(long long)lo, -(K + L));
}

// Sc[cid[v]] = S[v] for the vertices with out-edges, and the largest |S[v]|
// among them (as the bits of a non-negative double: their order is the
// values').  One atomic per workgroup at most, and only when its maximum
// passes the value already there (atomics on one address serialise: one per
// wave took 2.96 ms over 67 M vertices)
__global__ __launch_bounds__(kB) void k_beng_compact(const double* __restrict__ S, const uint32_t* __restrict__ cid,
                                                     uint32_t nV, double* __restrict__ Sc,
                                                     unsigned long long* __restrict__ smax_bits) {
  __shared__ unsigned long long wm[kB / 64];
  unsigned long long m = 0;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t k = cid[v];
    if (k < nV) {
      const double x = S[v];
      Sc[k] = x;
      const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(x));
      m = b > m ? b : m;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long y = __shfl_xor(m, o, 64);
    m = y > m ? y : m;
  }
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (unsigned k = 1; k < kB / 64; ++k) m = wm[k] > m ? wm[k] : m;
    if (m && m > __hip_atomic_load(smax_bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(smax_bits, m);
  }
}

// Row v's result: C[v] += its sum, or, with a fused update (f.X set), the
// update applied in place: X[v] = f.a + f.b (C[v] + sum), C[v] = 0.
__device__ inline void eng_row_done(double* __restrict__ C, const EngFuse& f, uint32_t v, double sum) {
  const double cv = C[v] + sum;
  if (f.X) {
    f.X[v] = f.a + f.b * cv;
    C[v] = 0.0;
  } else {
    C[v] = cv;
  }
}

// Accumulate slice s of bin b of the engine's layout (rows = the vertices with
// in-edges in vertex order): the two-word sums in LDS, then row v's result
// (eng_row_done; the bin's only slice) or the sums added into gacc (finished
// by k_beng_fin).  Rounds alternate between two register sets as in the FP32
// accumulate; lanes past the end add zero to the first slot's row.
template <int U>
__global__ __launch_bounds__(kAccumBlock) void k_beng_accum(
    const double* __restrict__ binval, const uint16_t* __restrict__ dloc, const uint32_t* __restrict__ sl_e0,
    const uint32_t* __restrict__ sl_bin, const uint32_t* __restrict__ sl_g, const uint32_t* __restrict__ bin_r0,
    const uint32_t* __restrict__ rowv, const unsigned long long* __restrict__ smax_bits, uint32_t max_in,
    double* __restrict__ C, EngFuse f, unsigned long long* __restrict__ gacc) {
  __shared__ unsigned long long ahi[kEngBinRows], alo[kEngBinRows];
  const uint32_t s = blockIdx.x;
  const uint32_t b = sl_bin[s];
  const uint32_t r0 = bin_r0[b], nr = bin_r0[b + 1] - r0;
  const uint32_t e0 = sl_e0[s], n = sl_e0[s + 1] - e0;
  const double* __restrict__ vp = binval + e0;
  const uint16_t* __restrict__ dp = dloc + e0;
  int K, L;
  eng_scales(smax_bits, max_in, K, L);
  constexpr uint32_t step = kAccumBlock * U;
  auto load = [&](double (&v)[U], uint32_t (&d)[U], uint32_t base) {
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const uint32_t i = base + j * kAccumBlock + threadIdx.x;
      const uint32_t ii = i < n ? i : 0u;
      v[j] = __builtin_nontemporal_load(vp + ii);
      d[j] = __builtin_nontemporal_load(dp + ii);
    }
  };
  auto add = [&](const double (&v)[U], const uint32_t (&d)[U], uint32_t base) {
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const uint32_t i = base + j * kAccumBlock + threadIdx.x;
      long long hi, lo;
      eng_split(i < n ? v[j] : 0.0, K, L, hi, lo);
      atomicAdd(&ahi[d[j]], (unsigned long long)hi);
      atomicAdd(&alo[d[j]], (unsigned long long)lo);
    }
  };
  double va[U], vb[U];
  uint32_t da[U], db[U];
  load(va, da, 0);  // in flight while the accumulators are cleared
  for (uint32_t i = threadIdx.x; i < nr; i += kAccumBlock) {
    ahi[i] = 0ULL;
    alo[i] = 0ULL;
  }
  __syncthreads();
  for (uint32_t base = 0;;) {
    load(vb, db, base + step);
    __builtin_amdgcn_sched_barrier(0);
    add(va, da, base);
    base += step;
    if (base >= n) break;
    load(va, da, base + step);
    __builtin_amdgcn_sched_barrier(0);
    add(vb, db, base);
    base += step;
    if (base >= n) break;
  }
  __syncthreads();
  const uint32_t g = sl_g[s];
  if (g == 0xffffffffu) {
    for (uint32_t i = threadIdx.x; i < nr; i += kAccumBlock)
      eng_row_done(C, f, rowv[r0 + i], eng_join(ahi[i], alo[i], K, L));
  } else {
    for (uint32_t i = threadIdx.x; i < nr; i += kAccumBlock) {
      if (ahi[i]) atomicAdd(&gacc[2 * (size_t)(g + i)], ahi[i]);
      if (alo[i]) atomicAdd(&gacc[2 * (size_t)(g + i) + 1], alo[i]);
    }
  }
}

// the rows of the engine layout's bins cut into slices (their two-word sums in
// gacc, cleared for the next superstep)
__global__ void k_beng_fin(const uint32_t* __restrict__ mb_r0, const uint32_t* __restrict__ mb_g, uint32_t nmb,
                           uint32_t nmrows, const uint32_t* __restrict__ rowv,
                           const unsigned long long* __restrict__ smax_bits, uint32_t max_in, double* __restrict__ C,
                           EngFuse f, unsigned long long* __restrict__ gacc) {
  int K, L;
  eng_scales(smax_bits, max_in, K, L);
  for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < nmrows; q += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = nmb;  // largest k with mb_g[k] <= q
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (mb_g[mid] <= q) lo = mid;
      else hi = mid;
    }
    const unsigned long long h = gacc[2 * q], l = gacc[2 * q + 1];
    gacc[2 * q] = 0ULL;
    gacc[2 * q + 1] = 0ULL;
    eng_row_done(C, f, rowv[mb_r0[lo] + (uint32_t)(q - mb_g[lo])], eng_join(h, l, K, L));
  }
}

// ---- layout build -----------------------------------------------------------

// rows of the vertices with in-edges: in-degree rank i dealt to bin b in snake
// order, row = bin_r0[b] + round
__global__ void k_bin_destrows(const uint32_t* __restrict__ did, uint32_t nV, uint32_t B,
                               const uint32_t* __restrict__ bin_r0, uint32_t* __restrict__ rowperm) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t i = did[v];
    if (i == 0xffffffffu) continue;
    const uint32_t k = i / B, p = i % B;
    const uint32_t b = (k & 1u) ? B - 1 - p : p;
    rowperm[v] = bin_r0[b] + k;
  }
}

// sort keys of the tail: vertices without in-edges by descending out-degree,
// the others last
__global__ void k_bin_tailkeys(const uint64_t* __restrict__ coff, const int32_t* __restrict__ outdeg, uint32_t nV,
                               uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV; v += (uint64_t)gridDim.x * blockDim.x) {
    key[v] = coff[v + 1] == coff[v] ? 0x7fffffffu - (uint32_t)outdeg[v] : 0xffffffffu;
    val[v] = (uint32_t)v;
  }
}

__global__ void k_bin_tailrows(const uint32_t* __restrict__ order, uint32_t ntail, uint32_t nd,
                               uint32_t* __restrict__ rowperm) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < ntail; j += (uint64_t)gridDim.x * blockDim.x)
    rowperm[order[j]] = nd + (uint32_t)j;
}

__global__ void k_bin_rowdata(const uint32_t* __restrict__ rowperm, const int32_t* __restrict__ outdeg, uint32_t nV,
                              uint32_t* __restrict__ rowinv, int32_t* __restrict__ outdeg_r,
                              uint32_t* __restrict__ cid_r) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = rowperm[v];
    rowinv[r] = (uint32_t)v;
    outdeg_r[r] = outdeg[v];
    cid_r[r] = outdeg[v] > 0 ? r : 0xffffffffu;
  }
}

__global__ void k_bin_rowbin(const uint32_t* __restrict__ bin_r0, uint32_t B, uint32_t nd,
                             uint32_t* __restrict__ rowbin) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < nd; r += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = B;  // largest b with bin_r0[b] <= r
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (bin_r0[mid] <= r) lo = mid;
      else hi = mid;
    }
    rowbin[r] = lo;
  }
}

// Sort key of CSR entry e (u -> v): (chunk, bin of v's row, v's row within
// the bin); value = (e, u's staging index within the chunk).  srcmap gives a
// vertex's staging index (the order the chunks cut), dstmap its row.  The row
// of e is found inside e's tile of the CSR's row tiles.
__global__ void k_bin_keys(const uint64_t* __restrict__ off, const uint32_t* __restrict__ idx,
                           const uint32_t* __restrict__ rb, uint32_t te, uint64_t E,
                           const uint32_t* __restrict__ srcmap, const uint32_t* __restrict__ dstmap,
                           const int32_t* __restrict__ outdeg, const uint32_t* __restrict__ row_ch,
                           const uint32_t* __restrict__ ch_r0, const uint32_t* __restrict__ rowbin,
                           const uint32_t* __restrict__ bin_r0, int bbits, uint64_t* __restrict__ key,
                           uint64_t* __restrict__ val) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E; e += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t t = e / te;
    uint32_t lo = rb[t] > 0 ? rb[t] - 1 : 0, hi = rb[t + 1];
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (off[mid] <= e) lo = mid;
      else hi = mid;
    }
    const uint32_t u = lo;
    const uint32_t ru = srcmap[u];
    const bool split = (uint32_t)outdeg[u] > kChunkEdges;
    const uint32_t w = row_ch[ru] + (split ? (uint32_t)((e - off[u]) / kChunkEdges) : 0u);
    const uint32_t sl = ru - ch_r0[w];
    const uint32_t rv = dstmap[idx[e]];
    const uint32_t b = rowbin[rv];
    key[e] = ((uint64_t)w << (bbits + kDstBits)) | ((uint64_t)b << kDstBits) | (rv - bin_r0[b]);
    val[e] = (e << 32) | sl;
  }
}

// each entry's word (source and flags) at its padded word position, the
// run-start and group-start bits (a group starts with a new destination, a new
// run or a 16-word row), and the first and one-past-last entry of every run
__global__ void k_bin_runs(const uint64_t* __restrict__ key, const uint64_t* __restrict__ val, uint64_t E, int bbits,
                           uint32_t B, const uint32_t* __restrict__ ch_e0, const uint32_t* __restrict__ ch_w0,
                           uint16_t* __restrict__ src16, unsigned long long* __restrict__ rbits,
                           unsigned long long* __restrict__ gbits, uint32_t* __restrict__ rstart,
                           uint32_t* __restrict__ rend) {
  const uint64_t bmask = (1ULL << bbits) - 1;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < E; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = key[i];
    const uint64_t wb = k >> kDstBits;
    const uint32_t b = (uint32_t)(wb & bmask);
    const uint64_t w = wb >> bbits;
    const uint32_t g = ch_w0[w] + ((uint32_t)i - ch_e0[w]);
    const bool run = i == 0 || (key[i - 1] >> kDstBits) != wb;
    const bool group = run || key[i - 1] != k || (g % kGroupLanes) == 0;
    src16[g] = (uint16_t)((val[i] & ((1u << kSrcBits) - 1)) | (run ? kWordRun : 0u) | (group ? kWordGroup : 0u));
    if (run) {
      rstart[w * B + b] = (uint32_t)i;
      atomicOr(&rbits[g >> 6], 1ULL << (g & 63));
    }
    if (group) atomicOr(&gbits[g >> 6], 1ULL << (g & 63));
    if (i + 1 == E || (key[i + 1] >> kDstBits) != wb) rend[w * B + b] = (uint32_t)(i + 1);
  }
}

// for each group's first entry: its destination at the group's bin-array slot;
// for every entry: its CSR position at its word.  bad[0] counts slots out of
// range, bad[1] accumulates the slots (their sum is checked on the host)
__global__ void k_bin_dloc(const uint64_t* __restrict__ key, const uint64_t* __restrict__ val, uint64_t E,
                           uint64_t Eb, int bbits, uint32_t B, const uint32_t* __restrict__ ch_e0,
                           const uint32_t* __restrict__ ch_w0, const uint32_t* __restrict__ runbase,
                           const unsigned long long* __restrict__ gbits, const uint32_t* __restrict__ rcnt,
                           uint16_t* __restrict__ dloc, uint32_t* __restrict__ wpos,
                           unsigned long long* __restrict__ bad) {
  const uint64_t bmask = (1ULL << bbits) - 1;
  unsigned long long nbad = 0, psum = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < E; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = key[i];
    const uint64_t wb = k >> kDstBits;
    const uint32_t b = (uint32_t)(wb & bmask);
    const uint64_t w = wb >> bbits;
    const uint32_t g = ch_w0[w] + ((uint32_t)i - ch_e0[w]);
    if (wpos) wpos[g] = (uint32_t)(val[i] >> 32);
    const unsigned long long gb = gbits[g >> 6];
    if (!((gb >> (g & 63)) & 1ULL)) continue;
    const uint32_t c = (rcnt[g >> 6] >> kRunBits) + (uint32_t)__popcll(gb & ((1ULL << (g & 63)) - 1));
    const uint32_t pos = c + runbase[w * B + b];
    if (pos < Eb) {
      dloc[pos] = (uint16_t)(k & ((1u << kDstBits) - 1));
      psum += pos;
    } else {
      ++nbad;
    }
  }
  block_add2(nbad, bad, psum, bad + 1);
}

// engine layout: flag[v] = v has in-edges; then rows in vertex order
__global__ void k_beng_flags(const uint64_t* __restrict__ coff, uint32_t nV, uint32_t* __restrict__ flag,
                             uint32_t* __restrict__ max_in) {
  uint32_t m = 0;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t d = coff[v + 1] - coff[v];
    flag[v] = d > 0 ? 1u : 0u;
    m = d > m ? (uint32_t)d : m;
  }
  m = wave_max_u32(m);
  if ((threadIdx.x & 63) == 0) atomicMax(max_in, m);
}

__global__ void k_beng_rows(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos, uint32_t nV,
                            uint32_t* __restrict__ dmap, uint32_t* __restrict__ rowv) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV; v += (uint64_t)gridDim.x * blockDim.x) {
    if (flag[v]) {
      dmap[v] = pos[v];
      rowv[pos[v]] = (uint32_t)v;
    } else {
      dmap[v] = 0xffffffffu;
    }
  }
}

// out-degree by compact source id
__global__ void k_beng_degc(const uint32_t* __restrict__ cid, const int32_t* __restrict__ outdeg, uint32_t nV,
                            int32_t* __restrict__ degc) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t k = cid[v];
    if (k < nV) degc[k] = outdeg[v];
  }
}

template <typename T>
void copy_to_dev(DevBuf& d, const std::vector<T>& h, hipStream_t s) {
  d.alloc(sizeof(T) * std::max<size_t>(h.size(), 1));
  if (!h.empty()) GRE_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
}

}  // namespace

namespace {

// rows of the PageRank handles' layout: the vertices with in-edges dealt to
// the bins by in-degree in snake order, then the others (with out-edges
// first, by descending out-degree); rowinv, out-degree and compact id by row
void build_rows_snake(gre_graph* g, XPart* xp, const BinPlan& bp, uint32_t B, hipStream_t s) {
  const uint32_t nV = g->nV, nd = bp.nd;
  xp->rowperm.alloc(4 * (size_t)nV);
  xp->rowinv.alloc(4 * (size_t)nV);
  xp->outdeg_r.alloc(4 * (size_t)nV);
  xp->cid_r.alloc(4 * (size_t)nV);
  hipLaunchKernelGGL(k_bin_destrows, dim3(gcap(nV)), dim3(kB), 0, s, g->did.as<uint32_t>(), nV, B,
                     bp.bin_r0.as<uint32_t>(), xp->rowperm.as<uint32_t>());
  GRE_HIP(hipGetLastError());
  if (nd < nV) {
    DevBuf k0, k1, v0, v1;
    k0.alloc(4 * (size_t)nV);
    k1.alloc(4 * (size_t)nV);
    v0.alloc(4 * (size_t)nV);
    v1.alloc(4 * (size_t)nV);
    hipLaunchKernelGGL(k_bin_tailkeys, dim3(gcap(nV)), dim3(kB), 0, s, g->coff.as<uint64_t>(),
                       g->outdeg.as<int32_t>(), nV, k0.as<uint32_t>(), v0.as<uint32_t>());
    GRE_HIP(hipGetLastError());
    size_t tb = 0;
    GRE_HIP(rocprim::radix_sort_pairs(nullptr, tb, k0.as<uint32_t>(), k1.as<uint32_t>(), v0.as<uint32_t>(),
                                      v1.as<uint32_t>(), (size_t)nV, 0, 32, s));
    void* tmp = scratch(0, tb);
    GRE_HIP(rocprim::radix_sort_pairs(tmp, tb, k0.as<uint32_t>(), k1.as<uint32_t>(), v0.as<uint32_t>(),
                                      v1.as<uint32_t>(), (size_t)nV, 0, 32, s));
    hipLaunchKernelGGL(k_bin_tailrows, dim3(gcap(nV - nd)), dim3(kB), 0, s, v1.as<uint32_t>(), nV - nd, nd,
                       xp->rowperm.as<uint32_t>());
    GRE_HIP(hipGetLastError());
    GRE_HIP(hipStreamSynchronize(s));
  }
  hipLaunchKernelGGL(k_bin_rowdata, dim3(gcap(nV)), dim3(kB), 0, s, xp->rowperm.as<uint32_t>(),
                     g->outdeg.as<int32_t>(), nV, xp->rowinv.as<uint32_t>(), xp->outdeg_r.as<int32_t>(),
                     xp->cid_r.as<uint32_t>());
  GRE_HIP(hipGetLastError());

}

// The binned layout.  eng = false: the PageRank handles' (bins of kBinRows
// rows dealt by in-degree in snake order, one row order for sources and
// destinations, ranks and contributions kept in it).  eng = true: the engine's
// FP64 sum pull (bins of kEngBinRows rows: the vertices with in-edges in
// vertex order, so each bin's results go to ascending vertices; sources staged
// in compact source order, hub-sorted).
std::shared_ptr<XPart> build_binpart(gre_graph* g, bool eng) {
  const auto mkey = std::make_pair((int)kBinKey, eng ? 1u : 0u);
  auto it = g->xpart.find(mkey);
  if (it != g->xpart.end()) return it->second;
  const uint64_t E = g->nE;
  const uint32_t nV = g->nV;
  GRE_CHECK(E > 0 && E < 0xffffffffULL, GRE_DEV_NULL, "the binned schedule needs 0 < edges < 2^32");
  graph_compact_targets(g);
  if (eng) graph_compact_sources(g);
  const uint32_t nd = g->ntgt;
  const uint32_t brows = eng ? (uint32_t)kEngBinRows : (uint32_t)kBinRows;
  const uint32_t B = (nd + brows - 1) / brows;
  GRE_CHECK(B >= 1 && B <= (uint32_t)kBinMaxBins, GRE_DEV_NULL,
            "the binned layout takes at most 4096 bins of vertices with in-edges");
  hipStream_t s = stream();
  auto xp = std::make_shared<XPart>();
  xp->binned = true;
  xp->permuted = !eng;
  xp->R = (int)kBinKey;
  BinPlan& bp = xp->bin;
  bp.E = E;
  bp.nd = nd;
  bp.B = B;
  bp.bin_rows = brows;

  std::vector<uint32_t> bin_r0(B + 1, 0);
  if (eng) {
    for (uint32_t b = 0; b < B; ++b) bin_r0[b + 1] = std::min(nd, (b + 1) * brows);
  } else {
    // B - rem bins get nd / B rows, rem get one more (the last, partial round
    // of the snake deal)
    const uint32_t full = nd / B, rem = nd % B;
    const bool even = (full % 2) == 0;
    for (uint32_t b = 0; b < B; ++b) {
      const bool extra = even ? b < rem : b >= B - rem;
      bin_r0[b + 1] = bin_r0[b] + full + (extra ? 1u : 0u);
    }
  }
  copy_to_dev(bp.bin_r0, bin_r0, s);

  // staging index and row of every vertex, and the staging order's degrees
  const uint32_t* srcmap = nullptr;
  const uint32_t* dstmap = nullptr;
  uint32_t nS = nV;
  std::vector<int32_t> deg;
  DevBuf dmap;
  if (eng) {
    DevBuf flag, pos, mx;
    flag.alloc(4 * (size_t)nV);
    pos.alloc(4 * (size_t)nV);
    mx.alloc(4);
    GRE_HIP(hipMemsetAsync(mx.p, 0, 4, s));
    hipLaunchKernelGGL(k_beng_flags, dim3(gcap(nV)), dim3(kB), 0, s, g->coff.as<uint64_t>(), nV, flag.as<uint32_t>(),
                       mx.as<uint32_t>());
    GRE_HIP(hipGetLastError());
    size_t tb = 0;
    GRE_HIP(rocprim::exclusive_scan(nullptr, tb, flag.as<uint32_t>(), pos.as<uint32_t>(), 0u, (size_t)nV,
                                    rocprim::plus<uint32_t>(), s));
    void* tmp = scratch(0, tb);
    GRE_HIP(rocprim::exclusive_scan(tmp, tb, flag.as<uint32_t>(), pos.as<uint32_t>(), 0u, (size_t)nV,
                                    rocprim::plus<uint32_t>(), s));
    dmap.alloc(4 * (size_t)nV);
    bp.rowv.alloc(4 * (size_t)std::max<uint32_t>(nd, 1));
    hipLaunchKernelGGL(k_beng_rows, dim3(gcap(nV)), dim3(kB), 0, s, flag.as<uint32_t>(), pos.as<uint32_t>(), nV,
                       dmap.as<uint32_t>(), bp.rowv.as<uint32_t>());
    GRE_HIP(hipGetLastError());
    nS = g->nsrc;
    DevBuf degc;
    degc.alloc(4 * (size_t)std::max<uint32_t>(nS, 1));
    hipLaunchKernelGGL(k_beng_degc, dim3(gcap(nV)), dim3(kB), 0, s, g->cid.as<uint32_t>(), g->outdeg.as<int32_t>(),
                       nV, degc.as<int32_t>());
    GRE_HIP(hipGetLastError());
    deg.resize(nS);
    GRE_HIP(hipMemcpyAsync(deg.data(), degc.p, 4 * (size_t)nS, hipMemcpyDeviceToHost, s));
    GRE_HIP(hipMemcpyAsync(&bp.max_in, mx.p, 4, hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
    srcmap = g->cid.as<uint32_t>();
    dstmap = dmap.as<uint32_t>();
  } else {
    build_rows_snake(g, xp.get(), bp, B, s);
    deg.resize(nV);
    GRE_HIP(hipMemcpyAsync(deg.data(), xp->outdeg_r.p, 4 * (size_t)nV, hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
    srcmap = xp->rowperm.as<uint32_t>();
    dstmap = xp->rowperm.as<uint32_t>();
  }

  // scatter chunks over consecutive staging indices
  const uint32_t crows = eng ? (uint32_t)kEngChunkRows : (uint32_t)kChunkRows;
  bp.chunk_rows = crows;
  std::vector<uint32_t> ch_r0, ch_nr, ch_e0(1, 0), row_ch(nS, 0);
  for (uint32_t r = 0; r < nS;) {
    const uint32_t d = (uint32_t)deg[r];
    if (d == 0) { ++r; continue; }
    if (d > kChunkEdges) {
      row_ch[r] = (uint32_t)ch_r0.size();
      for (uint32_t done = 0; done < d; done += kChunkEdges) {
        ch_r0.push_back(r);
        ch_nr.push_back(1);
        ch_e0.push_back(ch_e0.back() + std::min(kChunkEdges, d - done));
      }
      ++r;
      continue;
    }
    const uint32_t r0 = r, w = (uint32_t)ch_r0.size();
    uint32_t edges = 0;
    while (r < nS && r - r0 < crows) {
      const uint32_t dr = (uint32_t)deg[r];
      if (dr > kChunkEdges || edges + dr > kChunkEdges) break;
      edges += dr;
      row_ch[r] = w;
      ++r;
    }
    ch_r0.push_back(r0);
    ch_nr.push_back(r - r0);
    ch_e0.push_back(ch_e0.back() + edges);
  }
  const uint32_t nch = (uint32_t)ch_r0.size();
  GRE_CHECK(ch_e0.back() == E, GRE_DEV_FAILURE, "binned layout: chunk entries do not add up to the edges");
  bp.nch = nch;
  // the scatter's XCD stretches: XCD x runs chunks [xs[x], xs[x+1]) in order,
  // cut at equal shares of the entries (equal chunk counts gave the stretch
  // of the last, low-degree sources 0.4 of a share and the others 1.09)
  bp.xs.s[0] = 0;
  for (uint32_t x = 1, w = 0; x < 8; ++x) {
    const uint64_t want = (uint64_t)E * x / 8;
    while (w < nch && ch_e0[w] < want) ++w;
    bp.xs.s[x] = std::max(w, bp.xs.s[x - 1]);
  }
  bp.xs.s[8] = nch;
  bp.xs_max = 0;
  for (uint32_t x = 0; x < 8; ++x) bp.xs_max = std::max(bp.xs_max, bp.xs.s[x + 1] - bp.xs.s[x]);
  copy_to_dev(bp.ch_r0, ch_r0, s);
  copy_to_dev(bp.ch_nr, ch_nr, s);
  copy_to_dev(bp.ch_e0, ch_e0, s);
  // each chunk's words start on a 64-entry boundary
  std::vector<uint32_t> ch_w0(nch);
  uint64_t Ew = 0;
  for (uint32_t w = 0; w < nch; ++w) {
    ch_w0[w] = (uint32_t)Ew;
    Ew += align64(ch_e0[w + 1] - ch_e0[w]);
  }
  GRE_CHECK(Ew < 0xffffffffULL, GRE_DEV_NULL, "the binned schedule needs fewer than 2^32 padded entries");
  bp.Ew = Ew;
  copy_to_dev(bp.ch_w0, ch_w0, s);

  // entries sorted by (chunk, bin, destination row), ties in CSR order
  const int bbits = (int)bits_of(B);
  const int wbits = (int)bits_of((uint64_t)nch);
  const int kbits = wbits + bbits + kDstBits;
  GRE_CHECK(kbits <= 64, GRE_DEV_FAILURE, "binned layout: sort key too wide");
  DevBuf keys0, keys1, vals0, vals1;
  {
    DevBuf rowch_d, rowbin;
    copy_to_dev(rowch_d, row_ch, s);
    rowbin.alloc(4 * (size_t)std::max<uint32_t>(nd, 1));
    hipLaunchKernelGGL(k_bin_rowbin, dim3(gcap(nd)), dim3(kB), 0, s, bp.bin_r0.as<uint32_t>(), B, nd,
                       rowbin.as<uint32_t>());
    uint64_t nt = 0;
    const uint32_t* rb = graph_tiles(g, false, 2048, &nt);
    keys0.alloc(8 * E);
    vals0.alloc(8 * E);
    hipLaunchKernelGGL(k_bin_keys, dim3(gcap(E)), dim3(kB), 0, s, g->off.as<uint64_t>(), g->idx.as<uint32_t>(), rb,
                       2048u, E, srcmap, dstmap, g->outdeg.as<int32_t>(), rowch_d.as<uint32_t>(),
                       bp.ch_r0.as<uint32_t>(), rowbin.as<uint32_t>(), bp.bin_r0.as<uint32_t>(), bbits,
                       keys0.as<uint64_t>(), vals0.as<uint64_t>());
    GRE_HIP(hipGetLastError());
    GRE_HIP(hipStreamSynchronize(s));
  }
  keys1.alloc(8 * E);
  vals1.alloc(8 * E);
  {
    size_t tb = 0;
    GRE_HIP(rocprim::radix_sort_pairs(nullptr, tb, keys0.as<uint64_t>(), keys1.as<uint64_t>(), vals0.as<uint64_t>(),
                                      vals1.as<uint64_t>(), (size_t)E, 0, (unsigned)kbits, s));
    void* tmp = scratch(0, tb);
    GRE_HIP(rocprim::radix_sort_pairs(tmp, tb, keys0.as<uint64_t>(), keys1.as<uint64_t>(), vals0.as<uint64_t>(),
                                      vals1.as<uint64_t>(), (size_t)E, 0, (unsigned)kbits, s));
    GRE_HIP(hipStreamSynchronize(s));
  }
  keys0.release();
  vals0.release();

  // scatter sources, run-start and group-start bits, run bounds
  const size_t ntab = (size_t)nch * B;
  const size_t nblk = Ew / 64;
  bp.src16.alloc(2 * Ew);
  DevBuf rstart, rend, rbits, gbits;
  rstart.alloc(4 * ntab);
  rend.alloc(4 * ntab);
  rbits.alloc(8 * nblk);
  gbits.alloc(8 * nblk);
  GRE_HIP(hipMemsetAsync(rstart.p, 0, 4 * ntab, s));
  GRE_HIP(hipMemsetAsync(rend.p, 0, 4 * ntab, s));
  GRE_HIP(hipMemsetAsync(rbits.p, 0, 8 * nblk, s));
  GRE_HIP(hipMemsetAsync(gbits.p, 0, 8 * nblk, s));
  GRE_HIP(hipMemsetAsync(bp.src16.p, 0, 2 * Ew, s));
  hipLaunchKernelGGL(k_bin_runs, dim3(gcap(E)), dim3(kB), 0, s, keys1.as<uint64_t>(), vals1.as<uint64_t>(), E, bbits,
                     B, bp.ch_e0.as<uint32_t>(), bp.ch_w0.as<uint32_t>(), bp.src16.as<uint16_t>(),
                     rbits.as<unsigned long long>(), gbits.as<unsigned long long>(), rstart.as<uint32_t>(),
                     rend.as<uint32_t>());
  GRE_HIP(hipGetLastError());
  std::vector<uint32_t> hs(ntab), he(ntab);
  std::vector<uint64_t> hr(nblk), hg(nblk);
  GRE_HIP(hipMemcpyAsync(hs.data(), rstart.p, 4 * ntab, hipMemcpyDeviceToHost, s));
  GRE_HIP(hipMemcpyAsync(he.data(), rend.p, 4 * ntab, hipMemcpyDeviceToHost, s));
  GRE_HIP(hipMemcpyAsync(hr.data(), rbits.p, 8 * nblk, hipMemcpyDeviceToHost, s));
  GRE_HIP(hipMemcpyAsync(hg.data(), gbits.p, 8 * nblk, hipMemcpyDeviceToHost, s));
  GRE_HIP(hipStreamSynchronize(s));
  rstart.release();
  rend.release();
  rbits.release();
  // per 64-word block: the chunk's runs and groups before the block
  std::vector<uint32_t> rcnt(nblk, 0);
  for (uint32_t w = 0; w < nch; ++w) {
    const uint32_t n = ch_e0[w + 1] - ch_e0[w];
    const uint64_t k0 = ch_w0[w] / 64, k1 = k0 + align64(n) / 64;
    uint32_t runs = 0, groups = 0;
    for (uint64_t k = k0; k < k1; ++k) {
      const uint64_t r = hr[k], gb = hg[k];
      const uint32_t real = k + 1 == k1 && (n & 63) ? (n & 63) : 64u;
      uint64_t rows = 0;  // the row starts among the block's entries
      for (uint32_t q = 0; q < real; q += kGroupLanes) rows |= 1ULL << q;
      GRE_CHECK((gb & rows) == rows && (runs > 0 || (r & 1)), GRE_DEV_FAILURE,
                "binned layout: a row or chunk does not start a group");
      rcnt[k] = (runs + (uint32_t)(r & 1) - 1u) | (groups << kRunBits);
      runs += (uint32_t)__builtin_popcountll(r);
      groups += (uint32_t)__builtin_popcountll(gb);
    }
    GRE_CHECK(runs <= (uint32_t)kBinMaxBins, GRE_DEV_FAILURE, "binned layout: too many runs in a chunk");
  }
  copy_to_dev(bp.rcnt, rcnt, s);
  // groups of run k, and the chunk's groups before its first entry
  auto group_count = [&](uint64_t g0, uint64_t g1) {  // group starts in words [g0, g1)
    uint64_t c = 0;
    for (uint64_t k = g0 / 64; k * 64 < g1; ++k) {
      uint64_t m = hg[k];
      const uint64_t lo = std::max<uint64_t>(g0, k * 64) - k * 64, hi = std::min<uint64_t>(g1, k * 64 + 64) - k * 64;
      if (hi < 64) m &= (1ULL << hi) - 1;
      m &= ~0ULL << lo;
      c += (uint64_t)__builtin_popcountll(m);
    }
    return c;
  };
  std::vector<uint32_t> rn(ntab, 0), rc0(ntab, 0);
  for (uint32_t w = 0; w < nch; ++w)
    for (uint32_t b = 0; b < B; ++b) {
      const size_t k = (size_t)w * B + b;
      if (he[k] == hs[k]) continue;
      const uint64_t g0 = ch_w0[w] + (hs[k] - ch_e0[w]), g1 = ch_w0[w] + (he[k] - ch_e0[w]);
      rn[k] = (uint32_t)group_count(g0, g1);
      rc0[k] = (rcnt[g0 / 64] >> kRunBits) + (uint32_t)__builtin_popcountll(hg[g0 / 64] & ((1ULL << (g0 & 63)) - 1));
    }

  // bin-array slots: bin b holds its runs' groups in chunk order from
  // bin_e0[b], a 64-slot boundary; bin_n[b] slots
  std::vector<uint64_t> bin_n(B, 0), bin_e0(B + 1, 0);
  uint64_t tot = 0, ngrp = 0;
  for (uint32_t w = 0; w < nch; ++w)
    for (uint32_t b = 0; b < B; ++b) {
      const size_t k = (size_t)w * B + b;
      bin_n[b] += rn[k];
      tot += he[k] - hs[k];
    }
  for (uint32_t b = 0; b < B; ++b) {
    bin_e0[b + 1] = bin_e0[b] + align64(bin_n[b]);
    ngrp += bin_n[b];
  }
  GRE_CHECK(tot == E, GRE_DEV_FAILURE, "binned layout: runs do not add up to the edges");
  const uint64_t Eb = bin_e0[B];
  GRE_CHECK(Eb < 0xffffffffULL, GRE_DEV_NULL, "the binned schedule needs fewer than 2^32 - 1 padded slots");
  bp.Eb = Eb;
  bp.ngrp = ngrp;
  std::vector<uint32_t> runbase(ntab, 0);
  {
    std::vector<uint64_t> pos(bin_e0.begin(), bin_e0.end() - 1);
    for (uint32_t w = 0; w < nch; ++w)
      for (uint32_t b = 0; b < B; ++b) {
        const size_t k = (size_t)w * B + b;
        if (he[k] == hs[k]) continue;
        runbase[k] = (uint32_t)pos[b] - rc0[k];
        pos[b] += rn[k];
      }
  }
  // the dense table serves the build; the scatter keeps each chunk's non-empty runs
  DevBuf runbase_d;
  copy_to_dev(runbase_d, runbase, s);
  {
    std::vector<uint32_t> q0(nch + 1, 0), rbc;
    for (uint32_t w = 0; w < nch; ++w) {
      for (uint32_t b = 0; b < B; ++b) {
        const size_t k = (size_t)w * B + b;
        if (he[k] != hs[k]) rbc.push_back(runbase[k]);
      }
      q0[w + 1] = (uint32_t)rbc.size();
    }
    copy_to_dev(bp.ch_q0, q0, s);
    copy_to_dev(bp.rbase, rbc, s);
  }

  // accumulate slices; bins with more than kSliceEdges slots are cut (at
  // 64-slot boundaries); a bin's last slice runs to the next bin's start
  std::vector<uint32_t> sl_e0, sl_bin, sl_g, mb_r0, mb_g(1, 0);
  for (uint32_t b = 0; b < B; ++b) {
    const uint64_t n = bin_n[b];
    const uint64_t ns = std::max<uint64_t>(1, (n + kSliceEdges - 1) / kSliceEdges);
    uint32_t gslot = 0xffffffffu;
    if (ns > 1) {
      gslot = mb_g.back();
      mb_r0.push_back(bin_r0[b]);
      mb_g.push_back(mb_g.back() + (bin_r0[b + 1] - bin_r0[b]));
    }
    for (uint64_t j = 0; j < ns; ++j) {
      sl_e0.push_back((uint32_t)(bin_e0[b] + ((j * n / ns) & ~63ULL)));
      sl_bin.push_back(b);
      sl_g.push_back(gslot);
    }
  }
  sl_e0.push_back((uint32_t)Eb);
  bp.nsl = (uint32_t)sl_bin.size();
  bp.nmb = (uint32_t)mb_r0.size();
  bp.nmrows = mb_g.back();
  copy_to_dev(bp.sl_e0, sl_e0, s);
  copy_to_dev(bp.sl_bin, sl_bin, s);
  copy_to_dev(bp.sl_g, sl_g, s);
  copy_to_dev(bp.mb_r0, mb_r0, s);
  copy_to_dev(bp.mb_g, mb_g, s);

  // destination of every group at its slot
  bp.dloc.alloc(2 * Eb);
  {
    DevBuf bad;
    bad.alloc(16);
    GRE_HIP(hipMemsetAsync(bad.p, 0, 16, s));
    GRE_HIP(hipMemsetAsync(bp.dloc.p, 0, 2 * Eb, s));
    if (!eng) {  // the engine's programs carry no edge weights
      bp.wpos.alloc(4 * Ew);
      GRE_HIP(hipMemsetAsync(bp.wpos.p, 0, 4 * Ew, s));
    }
    hipLaunchKernelGGL(k_bin_dloc, dim3(gcap(E)), dim3(kB), 0, s, keys1.as<uint64_t>(), vals1.as<uint64_t>(), E, Eb,
                       bbits, B, bp.ch_e0.as<uint32_t>(), bp.ch_w0.as<uint32_t>(), runbase_d.as<uint32_t>(),
                       gbits.as<unsigned long long>(), bp.rcnt.as<uint32_t>(), bp.dloc.as<uint16_t>(),
                       eng ? nullptr : bp.wpos.as<uint32_t>(), bad.as<unsigned long long>());
    GRE_HIP(hipGetLastError());
    unsigned long long h[2] = {0, 0};
    GRE_HIP(hipMemcpyAsync(h, bad.p, 16, hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
    // the slots of bin b are bin_e0[b] .. bin_e0[b] + bin_n[b] - 1
    unsigned long long want = 0;
    for (uint32_t b = 0; b < B; ++b) {
      const unsigned long long a = bin_e0[b], n = bin_n[b];
      want += n * a + (n % 2 ? n * ((n - 1) / 2) : (n / 2) * (n - 1));
    }
    GRE_CHECK(h[0] == 0 && h[1] == want, GRE_DEV_FAILURE, "binned layout: group slots are not a permutation");
  }
  g->xpart[mkey] = xp;
  return xp;
}

}  // namespace

std::shared_ptr<XPart> graph_binpart(gre_graph* g) { return build_binpart(g, false); }
std::shared_ptr<XPart> graph_binpart_engine(gre_graph* g) { return build_binpart(g, true); }

void BinBuffers::make(const XPart& xp, uint32_t nV, hipStream_t s) {
  const BinPlan& bp = xp.bin;
  // the gaps between bins are never scattered into and must hold zero
  // The scatter's writes go to every bin at once (~1,700 fronts 2 MiB apart at
  // R-MAT scale 26), so they need a translation for every bin's page.  A start
  // on a large power-of-two boundary lets the page tables map the array with
  // fragments larger than 2 MiB: at scale 26 the scatter took 1.32-1.45 ms with
  // the array aligned to 256 MiB or 1 GiB, and 1.69-1.85 ms at hipMalloc's
  // 2-4 MiB alignment.
  const size_t nb = 4 * (size_t)bp.Eb;
  size_t al = 0;
  if (nb >= (64u << 20)) {
    al = 64u << 20;
    while (al < (1u << 30) && 2 * al <= nb) al *= 2;
  }
  binval.alloc(nb, al);
  GRE_HIP(hipMemsetAsync(binval.p, 0, nb, s));
  gacc.alloc(8 * (size_t)std::max<uint32_t>(bp.nmrows, 1));
  GRE_HIP(hipMemsetAsync(gacc.p, 0, 8 * (size_t)std::max<uint32_t>(bp.nmrows, 1), s));
  tail_blocks = (uint32_t)(((uint64_t)(nV - bp.nd) + kTailRows - 1) / kTailRows);
  part.alloc(2 * sizeof(double) * ((size_t)bp.nsl + tail_blocks));
  fin_grid = gcap(std::max<uint32_t>(bp.nmrows, 1));
  part2.alloc(2 * sizeof(double) * fin_grid);
  part3.alloc(2 * sizeof(double) * kRedBlocks);
  GRE_HIP(hipStreamSynchronize(s));
  for_xp = &xp;
}

namespace {
template <bool R, bool W>
void bin_iter_t(const XPart& xp, const BinBuffers& bb, const BinIter& a, uint32_t nV, hipStream_t s) {
  const BinPlan& bp = xp.bin;
  using Up = BinUpdate<R, W>;
  Up up{a.contrib_out, a.x, a.p_r, a.outdeg_r, a.scal_in, (float)a.damp, (float)(1.0 - a.damp), (float)a.invn};
  {
    ProfScope ps("bin_scatter");
    auto kern = bp.Eb >= (1ULL << 30) ? k_bin_scatter<kScatterU, W, true> : k_bin_scatter<kScatterU, W, false>;
    hipLaunchKernelGGL(kern, dim3(8 * bp.xs_max), dim3(kScatterBlock), 0, s, a.contrib_in,
                       bp.ch_r0.as<uint32_t>(), bp.ch_nr.as<uint32_t>(), bp.ch_e0.as<uint32_t>(),
                       bp.ch_w0.as<uint32_t>(), bp.ch_q0.as<uint32_t>(), bp.src16.as<uint16_t>(),
                       bp.rcnt.as<uint32_t>(), bp.rbase.as<uint32_t>(), a.wbin, bp.xs,
                       (uint32_t)bp.Eb,
                       bb.binval.as<float>(), (float*)nullptr);
  }
  {
    ProfScope ps("bin_accum");
    hipLaunchKernelGGL((k_bin_accum<Up, kAccumU>), dim3(bp.nsl + bb.tail_blocks), dim3(kAccumBlock), 0, s,
                       bb.binval.as<float>(), bp.dloc.as<uint16_t>(), bp.sl_e0.as<uint32_t>(),
                       bp.sl_bin.as<uint32_t>(), bp.sl_g.as<uint32_t>(), bp.bin_r0.as<uint32_t>(), bp.nsl, bp.nd,
                       nV, up, bb.gacc.as<unsigned long long>(), bb.part.as<double>());
  }
  {
    ProfScope ps("bin_finish");
    hipLaunchKernelGGL((k_bin_fin<Up>), dim3(bb.fin_grid), dim3(kB), 0, s, bp.mb_r0.as<uint32_t>(),
                       bp.mb_g.as<uint32_t>(), bp.nmb, bp.nmrows, up, bb.gacc.as<unsigned long long>(),
                       bb.part2.as<double>());
  }
  reduce_pairs(bb.part.as<double>(), bp.nsl + bb.tail_blocks, bb.part2.as<double>(), bb.fin_grid, a.scal_out,
               bb.part3.as<double>(), s);
  GRE_HIP(hipGetLastError());
}
}  // namespace

void EngBinBuffers::make(const XPart& xp, uint32_t nsrc, hipStream_t s) {
  const BinPlan& bp = xp.bin;
  const size_t nb = 8 * (size_t)bp.Eb;
  size_t al = 0;  // aligned as the FP32 bin array
  if (nb >= (64u << 20)) {
    al = 64u << 20;
    while (al < (1u << 30) && 2 * al <= nb) al *= 2;
  }
  binval.alloc(nb, al);
  GRE_HIP(hipMemsetAsync(binval.p, 0, nb, s));
  trash.alloc(sizeof(double) * (size_t)kTrashBlocks * (kScatterBlock / 64) * (2 * kEngScatterU * 64));
  gacc.alloc(16 * (size_t)std::max<uint32_t>(bp.nmrows, 1));
  GRE_HIP(hipMemsetAsync(gacc.p, 0, 16 * (size_t)std::max<uint32_t>(bp.nmrows, 1), s));
  Sc.alloc(sizeof(double) * (size_t)std::max<uint32_t>(nsrc, 1));
  smax.alloc(8);
  GRE_HIP(hipStreamSynchronize(s));
  for_xp = &xp;
}

void bin_engine_sum(const XPart& xp, const EngBinBuffers& bb, const double* S, const uint32_t* cid, uint32_t nV,
                    bool fresh_s, double* C, const EngFuse& f, hipStream_t s) {
  const BinPlan& bp = xp.bin;
  if (fresh_s) {
    GRE_HIP(hipMemsetAsync(bb.smax.p, 0, 8, s));
    hipLaunchKernelGGL(k_beng_compact, dim3(gcap(nV)), dim3(kB), 0, s, S, cid, nV, bb.Sc.as<double>(),
                       bb.smax.as<unsigned long long>());
    GRE_HIP(hipGetLastError());
  }
  {
    ProfScope ps("beng_scatter");
    hipLaunchKernelGGL((k_bin_scatter<kEngScatterU, false, false, double, kEngChunkRows>), dim3(8 * bp.xs_max),
                       dim3(kScatterBlock), 0, s,
                       (const double*)bb.Sc.as<double>(), bp.ch_r0.as<uint32_t>(), bp.ch_nr.as<uint32_t>(),
                       bp.ch_e0.as<uint32_t>(), bp.ch_w0.as<uint32_t>(), bp.ch_q0.as<uint32_t>(),
                       bp.src16.as<uint16_t>(), bp.rcnt.as<uint32_t>(), bp.rbase.as<uint32_t>(), (const float*)nullptr,
                       bp.xs, (uint32_t)bp.Eb, bb.binval.as<double>(), bb.trash.as<double>());
    GRE_HIP(hipGetLastError());
  }
  {
    ProfScope ps("beng_accum");
    hipLaunchKernelGGL((k_beng_accum<kAccumU>), dim3(bp.nsl), dim3(kAccumBlock), 0, s, bb.binval.as<double>(),
                       bp.dloc.as<uint16_t>(), bp.sl_e0.as<uint32_t>(), bp.sl_bin.as<uint32_t>(),
                       bp.sl_g.as<uint32_t>(), bp.bin_r0.as<uint32_t>(), bp.rowv.as<uint32_t>(),
                       bb.smax.as<unsigned long long>(), bp.max_in, C, f, bb.gacc.as<unsigned long long>());
    GRE_HIP(hipGetLastError());
  }
  if (bp.nmrows) {
    ProfScope ps("beng_finish");
    hipLaunchKernelGGL(k_beng_fin, dim3(gcap(bp.nmrows)), dim3(kB), 0, s, bp.mb_r0.as<uint32_t>(),
                       bp.mb_g.as<uint32_t>(), bp.nmb, bp.nmrows, bp.rowv.as<uint32_t>(),
                       bb.smax.as<unsigned long long>(), bp.max_in, C, f, bb.gacc.as<unsigned long long>());
    GRE_HIP(hipGetLastError());
  }
}

void bin_forward_iter(const XPart& xp, const BinBuffers& bb, const BinIter& a, uint32_t nV, hipStream_t s) {
  if (a.wbin) {
    if (a.resid) bin_iter_t<true, true>(xp, bb, a, nV, s);
    else bin_iter_t<false, true>(xp, bb, a, nV, s);
  } else {
    if (a.resid) bin_iter_t<true, false>(xp, bb, a, nV, s);
    else bin_iter_t<false, false>(xp, bb, a, nV, s);
  }
}

// ---- the binned adjoint -------------------------------------------------------

// *mx = the larger of *mx and the bits of max |x[i]| (the bits of non-negative
// floats order as their values); ABS = false: of max x[i].  One atomic per
// workgroup, on a grid of at most kMaxGrid (atomics on one address serialise:
// one per wave over 67 M values took 2.9 ms)
constexpr unsigned kMaxGrid = 1024;
template <bool ABS>
__global__ __launch_bounds__(kB) void k_bin_max(const uint32_t* __restrict__ x, uint64_t n,
                                                uint32_t* __restrict__ mx) {
  __shared__ uint32_t wm[kB / 64];
  uint32_t m = 0;
  const uint64_t st = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  // eight loads in flight per thread (one at a time read 268 MB at 2 TB/s)
  for (; i + 7 * st < n; i += 8 * st) {
    uint32_t b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) b[u] = x[i + u * st];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t v = ABS ? b[u] & 0x7fffffffu : b[u];
      m = v > m ? v : m;
    }
  }
  for (; i < n; i += st) {
    const uint32_t v = ABS ? x[i] & 0x7fffffffu : x[i];
    m = v > m ? v : m;
  }
  m = wave_max_u32(m);
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (unsigned k = 1; k < kB / 64; ++k) m = wm[k] > m ? wm[k] : m;
    if (m) atomicMax(mx, m);
  }
}

__global__ void k_bin_zero_u32(uint32_t* __restrict__ p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

void bin_zero_u32(uint32_t* p, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_zero_u32, dim3(gcap(n)), dim3(kB), 0, s, p, n);
  GRE_HIP(hipGetLastError());
}
void bin_absmax(const float* x, uint64_t n, uint32_t* mx, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_max<true>, dim3(std::min(gcap(n), kMaxGrid)), dim3(kB), 0, s,
                     reinterpret_cast<const uint32_t*>(x), n, mx);
  GRE_HIP(hipGetLastError());
}
void bin_maxu(const uint32_t* x, uint64_t n, uint32_t* mx, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_max<false>, dim3(std::min(gcap(n), kMaxGrid)), dim3(kB), 0, s, x, n, mx);
  GRE_HIP(hipGetLastError());
}

namespace {
template <bool W>
void bin_adj_t(const XPart& xp, const BinBuffers& bb, const BinAdjIter& a, uint32_t nV, hipStream_t s) {
  const BinPlan& bp = xp.bin;
  using Up = BinAdjUpdate<W>;
  Up up{a.lout, a.wl, a.lin, a.c_r, a.p_r, a.deg_r, a.scal_in, a.mx_in, a.F, (float)a.damp, (float)a.invn};
  {
    ProfScope ps("bin_adj_scatter");
    auto kern = bp.Eb >= (1ULL << 30) ? k_bin_scatter<kScatterU, W, true> : k_bin_scatter<kScatterU, W, false>;
    hipLaunchKernelGGL(kern, dim3(8 * bp.xs_max), dim3(kScatterBlock), 0, s, a.lin, bp.ch_r0.as<uint32_t>(),
                       bp.ch_nr.as<uint32_t>(), bp.ch_e0.as<uint32_t>(), bp.ch_w0.as<uint32_t>(),
                       bp.ch_q0.as<uint32_t>(), bp.src16.as<uint16_t>(), bp.rcnt.as<uint32_t>(),
                       bp.rbase.as<uint32_t>(), a.wbin, bp.xs, (uint32_t)bp.Eb, bb.binval.as<float>(),
                       (float*)nullptr);
  }
  {
    ProfScope ps("bin_adj_accum");
    hipLaunchKernelGGL((k_bin_accum<Up, kAccumU>), dim3(bp.nsl + bb.tail_blocks), dim3(kAccumBlock), 0, s,
                       bb.binval.as<float>(), bp.dloc.as<uint16_t>(), bp.sl_e0.as<uint32_t>(),
                       bp.sl_bin.as<uint32_t>(), bp.sl_g.as<uint32_t>(), bp.bin_r0.as<uint32_t>(), bp.nsl, bp.nd,
                       nV, up, bb.gacc.as<unsigned long long>(), bb.part.as<double>());
  }
  {
    ProfScope ps("bin_adj_finish");
    hipLaunchKernelGGL((k_bin_fin<Up>), dim3(bb.fin_grid), dim3(kB), 0, s, bp.mb_r0.as<uint32_t>(),
                       bp.mb_g.as<uint32_t>(), bp.nmb, bp.nmrows, up, bb.gacc.as<unsigned long long>(),
                       bb.part2.as<double>());
  }
  reduce_pairs(bb.part.as<double>(), bp.nsl + bb.tail_blocks, bb.part2.as<double>(), bb.fin_grid, a.scal_out,
               bb.part3.as<double>(), s);
  GRE_HIP(hipGetLastError());
  bin_absmax(a.lout, nV, a.mx_out, s);
}
}  // namespace

void bin_adjoint_iter(const XPart& xp, const BinBuffers& bb, const BinAdjIter& a, uint32_t nV, hipStream_t s) {
  if (a.wbin) bin_adj_t<true>(xp, bb, a, nV, s);
  else bin_adj_t<false>(xp, bb, a, nV, s);
}

}  // namespace gre
