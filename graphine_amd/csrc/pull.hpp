// Tile-segmented pull: the combine of GRE's superstep done as an in-HBM
// segmented reduction over a compressed adjacency (CSC for a pull along
// out-edges, CSR for the transposed pass).
//
// The edge array is cut into fixed tiles of TE = 256*IPT entries; one
// 256-thread block owns a tile.  Stage 1 streams the tile's neighbour ids
// (coalesced) and gathers the neighbour values into LDS.  Stage 2 walks the
// rows that start in the tile (thread per row; rows longer than 32 in-tile
// entries go to a wave each).  A row that started in an earlier tile leaves a
// partial P[t]; a row that runs past the tile's end leaves O[t]; both are
// summed in tile order by a fix-up pass, so results are deterministic.
// Empty rows are finalised with a zero sum by the tile their offset lands in
// (the last tile also takes the trailing empty rows).
#pragma once
#include "graph.hpp"

#include <type_traits>
#include <utility>

namespace gre {

constexpr int kPullBlock = 256;
constexpr int kLongRow = 32;

// Fixed-order block reduction of two doubles; thread 0 gets the totals.
__device__ inline void block_sum2(double& a, double& b, double* sh /*[2*4]*/) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) { sh[w] = a; sh[4 + w] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = sh[0] + sh[1] + sh[2] + sh[3];
    b = sh[4] + sh[5] + sh[6] + sh[7];
  }
}

// Mode concept:
//   using T = ...;
//   __device__ T gather(uint64_t e, uint32_t nbr) const;
//   __device__ void finalize(uint32_t row, T sum, double& acc0, double& acc1) const;
// and optionally, for the row's own data to be loaded ahead with the tile:
//   __device__ Pre pre(uint32_t row) const;
//   __device__ void finalize_pre(uint32_t row, T sum, const Pre&, double& acc0, double& acc1) const;
template <class M, class = void>
struct HasPre : std::false_type {};
template <class M>
struct HasPre<M, std::void_t<decltype(std::declval<const M&>().pre(0u))>> : std::true_type {};
struct NoPre {};
template <class Mode>
__device__ __forceinline__ auto row_pre(const Mode& m, uint32_t r) {
  if constexpr (HasPre<Mode>::value) return m.pre(r);
  else return NoPre{};
}
template <class Mode, class P>
__device__ __forceinline__ void row_fin(const Mode& m, uint32_t r, typename Mode::T s, const P& q, double& a0,
                                        double& a1) {
  if constexpr (HasPre<Mode>::value) m.finalize_pre(r, s, q, a0, a1);
  else m.finalize(r, s, a0, a1);
}

// Fixed-order sum of two doubles over one 256-thread group (gid-th group of the
// block); tid 0 of the group gets the totals.  red holds 8 doubles per group.
// Every thread of the block must call it (it holds a block barrier).
__device__ inline void group_sum2(double& a, double& b, double* red, int tid) {
  a = wave_sum(a);
  b = wave_sum(b);
  const int w = tid >> 6, l = tid & 63;
  if (l == 0) { red[w] = a; red[4 + w] = b; }
  __syncthreads();
  if (tid == 0) {
    a = red[0] + red[1] + red[2] + red[3];
    b = red[4] + red[5] + red[6] + red[7];
  }
}

// Rows of a tile: thread j takes row rb - 1 + j, j in [0, re - rb + 1); row
// rb - 1 only when it continues into the tile.  The first Q rows of each
// thread have their offsets (and the mode's pre data) loaded together with the
// tile's gathers, so the tile waits for one memory latency less per step; the
// rest are loaded where they are walked.  Q trades that against registers.
template <class Mode, int Q, class Off>
struct TileRows {
  using Pre = decltype(row_pre(std::declval<const Mode&>(), 0u));
  Off ob[Q], oe[Q];
  Pre pre[Q];
  // issued without waiting: lanes past the rows load row 0
  __device__ __forceinline__ void load(const Off* __restrict__ off, uint32_t rb, uint32_t re, const Mode& m,
                                       int tid) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int64_t r = (int64_t)rb - 1 + tid + q * kPullBlock;
      const uint32_t rs = (r >= 0 && r < (int64_t)re) ? (uint32_t)r : 0u;
      ob[q] = off[rs];
      oe[q] = off[rs + 1];
      pre[q] = row_pre(m, rs);
    }
  }
};

// Stages 2 and 3 of tile t (entries e0..e1, n of them) whose values vals[0..n)
// are in LDS and *nlong = 0, both visible to the whole group; rows as above.
template <int IPT, class Mode, int Q, class Off>
__device__ __forceinline__ void tile_rows(
    uint64_t t, Off e0, Off e1, int n, const Off* __restrict__ off, uint32_t rb, uint32_t re,
    const TileRows<Mode, Q, Off>& tr, typename Mode::T* __restrict__ carryP, typename Mode::T* __restrict__ carryO,
    double* __restrict__ part, const Mode& m, const typename Mode::T* vals, uint32_t* longj, int* nlong, double* red,
    int tid) {
  using T = typename Mode::T;
  const uint32_t nrows = re - rb + 1;
  // row rb - 1 is thread 0's first row: tr.oe[0] = off[rb] there
  const bool has_rc = rb > 0 && n > 0 && tr.oe[0] > e0;
  double acc0 = 0.0, acc1 = 0.0;
  auto row = [&](uint32_t j, Off ob, Off oe, const typename TileRows<Mode, Q, Off>::Pre& q) {
    const uint32_t r = rb - 1 + j;
    const Off sb = ob > e0 ? ob : e0;
    const Off se = oe < e1 ? oe : e1;
    const int len = se > sb ? (int)(se - sb) : 0;
    if (len > kLongRow) {
      longj[atomicAdd(nlong, 1)] = j;
      return;
    }
    T s = (T)0;
    const int base = (int)(sb - e0);
    for (int i = 0; i < len; ++i) s += vals[base + i];
    if (ob < e0) carryP[t] = s;
    else if (oe > e1) carryO[t] = s;
    else row_fin(m, r, s, q, acc0, acc1);
  };
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const uint32_t j = (uint32_t)(tid + q * kPullBlock);
    if (j < nrows && (j > 0 || has_rc)) row(j, tr.ob[q], tr.oe[q], tr.pre[q]);
  }
  for (uint32_t j = (uint32_t)(tid + Q * kPullBlock); j < nrows; j += kPullBlock) {
    const uint32_t r = rb - 1 + j;
    row(j, off[r], off[r + 1], row_pre(m, r));
  }
  __syncthreads();

  // stage 3: long rows, one wave each
  const int w = tid >> 6, l = tid & 63;
  const int nl = *nlong;
  for (int q = w; q < nl; q += kPullBlock / 64) {
    const uint32_t r = rb - 1 + longj[q];
    const Off ob = off[r], oe = off[r + 1];
    const Off sb = ob > e0 ? ob : e0;
    const Off se = oe < e1 ? oe : e1;
    const int base = (int)(sb - e0), len = (int)(se - sb);
    T s = (T)0;
    for (int i = l; i < len; i += 64) s += vals[base + i];
    s = wave_sum(s);
    if (l == 0) {
      if (ob < e0) carryP[t] = s;
      else if (oe > e1) carryO[t] = s;
      else m.finalize(r, s, acc0, acc1);
    }
  }

  group_sum2(acc0, acc1, red, tid);
  if (tid == 0) { part[2 * t] = acc0; part[2 * t + 1] = acc1; }
}

// One tile of TE = 256*IPT entries, processed by one 256-thread group (tid in
// [0, 256)).  vals/longj/nlong/red are the group's LDS.  gat(e, u) gives the
// value gathered for entry e with neighbour u.  Off is the row offset type:
// uint32 where every offset fits (fewer bytes and 32-bit row arithmetic).
template <int IPT, bool NT, class Mode, class Gather, class Id, class Off>
__device__ __forceinline__ void pull_tile_body(
    uint64_t t, const Off* __restrict__ off, const Id* __restrict__ idx, const uint32_t* __restrict__ tile_rb,
    uint64_t E, uint32_t V, uint64_t ntiles, typename Mode::T* __restrict__ carryP,
    typename Mode::T* __restrict__ carryO, double* __restrict__ part, const Mode& m, const Gather& gat,
    typename Mode::T* vals, uint32_t* longj, int* nlong, double* red, int tid) {
  constexpr int TE = kPullBlock * IPT;
  const uint64_t e0l = t * (uint64_t)TE;
  const Off e0 = (Off)e0l;
  const Off e1 = (Off)(e0l + TE < E ? e0l + TE : E);
  const int n = e1 > e0 ? (int)(e1 - e0) : 0;
  const uint32_t rb = tile_rb[t];
  const uint32_t re = (t + 1 == ntiles) ? V : tile_rb[t + 1];

  // stage 1: neighbour ids (coalesced; NT = streamed with a non-temporal hint
  // so they do not push the gathered values out of the caches), then gathers
  uint32_t nb[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int i = j * kPullBlock + tid;
    if (NT) nb[j] = i < n ? (uint32_t)__builtin_nontemporal_load(idx + e0 + i) : 0u;
    else nb[j] = i < n ? (uint32_t)idx[e0 + i] : 0u;
  }
  // the rows' offsets are loaded in the same round as the gathers (issued
  // first: a load still in flight on only one side of the branch below would
  // make the code after it wait for every load)
  // rows loaded ahead per thread: 256*Q covers the ~TE/3.3 rows of a segment
  // plan's tile on R-MAT graphs; modes with pre data keep fewer in registers
  constexpr int Q = HasPre<Mode>::value ? (IPT >= 8 ? 2 : 1) : (IPT >= 8 ? 4 : 2);
  TileRows<Mode, Q, Off> tr;
  tr.load(off, rb, re, m, tid);
  // every gather is issued before any is used: a lane past the tile's end
  // reads entry e0 with neighbour 0 (nb = 0 there), a valid element, and
  // stores it to a slot of vals that no row reads (i < TE always).  Guarded
  // gathers or stores would make each load wait for the one before it.
  if (n > 0) {
    typename Mode::T gv[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const int i = j * kPullBlock + tid;
      gv[j] = gat(e0l + (i < n ? (uint64_t)i : 0ull), nb[j]);
    }
#pragma unroll
    for (int j = 0; j < IPT; ++j) vals[j * kPullBlock + tid] = gv[j];
  }
  if (tid == 0) *nlong = 0;
  __syncthreads();
  tile_rows<IPT>(t, e0, e1, n, off, rb, re, tr, carryP, carryO, part, m, vals, longj, nlong, red, tid);
}

template <class Mode>
struct ModeGather {
  const Mode& m;
  __device__ typename Mode::T operator()(uint64_t e, uint32_t u) const { return m.gather(e, u); }
};

template <int IPT, bool NT, class Mode, class Off>
__global__ __launch_bounds__(kPullBlock) void k_pull_tile(
    const Off* __restrict__ off, const uint32_t* __restrict__ idx,
    const uint32_t* __restrict__ tile_rb, uint64_t E, uint32_t V, uint64_t ntiles,
    typename Mode::T* __restrict__ carryP, typename Mode::T* __restrict__ carryO,
    double* __restrict__ part, const uint32_t* __restrict__ xmap, Mode m) {
  using T = typename Mode::T;
  constexpr int TE = kPullBlock * IPT;
  __shared__ T vals[TE];
  __shared__ uint32_t longj[TE / kLongRow + 2];
  __shared__ int nlong;
  __shared__ double red[8];

  // xmap: workgroup b runs tile xmap[c] + b / 8 of class c = b % 8, so every
  // tile of a class runs on the same XCD (workgroups go to XCD b % 8)
  uint64_t t = blockIdx.x;
  if (xmap) {
    const uint32_t c = blockIdx.x & 7u;
    t = (uint64_t)xmap[c] + (blockIdx.x >> 3);
    if (t >= xmap[c + 1]) return;
  }
  pull_tile_body<IPT, NT>(t, off, idx, tile_rb, E, V, ntiles, carryP, carryO, part, m, ModeGather<Mode>{m}, vals,
                          longj, &nlong, red, (int)threadIdx.x);
}

// The per-row pass over a partition's segment sums (rows = vertices, entries =
// each row's segment ids).  A tile's segments form one run of consecutive ids
// per (class, range) key; the tile first copies those runs from the mode's
// gather array into LDS in key order (coalesced: stage entry i belongs to the
// run whose start is the last set bit of the tile's run-start bitmap at or
// before i, and holds segment base[rank] + i), then pulls with each entry's
// value read from its LDS slot.  The entries, and so the order of every row's
// sum, are those of the plain tile pull over the segment ids.  The loads of
// the tile are issued in two rounds: (slots, run bitmap and bases, row
// bounds), then (staged sums, row offsets and data), so a tile waits for two
// memory latencies.
template <int IPT, class Mode, class Off>
__global__ __launch_bounds__(kPullBlock) void k_pull_staged(
    const Off* __restrict__ off, const uint16_t* __restrict__ slot, const uint32_t* __restrict__ tile_rb,
    uint64_t E, uint32_t V, uint64_t ntiles, const uint2* __restrict__ words, const uint32_t* __restrict__ base,
    int K, typename Mode::T* __restrict__ carryP, typename Mode::T* __restrict__ carryO, double* __restrict__ part,
    Mode m) {
  using T = typename Mode::T;
  constexpr int TE = kPullBlock * IPT;
  constexpr int W = TE / 32;
  static_assert((TE & (TE - 1)) == 0, "tile entries must be a power of two");
  __shared__ T stage[TE];
  __shared__ T vals[TE];
  __shared__ uint2 sW[W];
  __shared__ uint32_t sD[kMaxKeys];
  __shared__ uint32_t longj[TE / kLongRow + 2];
  __shared__ int nlong;
  __shared__ double red[8];
  // workgroup b runs on XCD b % 8; within each group of 64 tiles XCD x takes
  // tiles 8x .. 8x+7, so the consecutive tiles that fetch the two ends of the
  // same runs' lines share one L2 (a bijection; the last partial group keeps
  // its order)
  const uint64_t bid = blockIdx.x;
  const uint64_t t = bid < (ntiles & ~(uint64_t)63) ? (bid & ~(uint64_t)63) + ((bid & 7) << 3) + ((bid & 63) >> 3) : bid;
  const int tid = (int)threadIdx.x;
  const uint64_t e0l = t * (uint64_t)TE;
  const Off e0 = (Off)e0l;
  const Off e1 = (Off)(e0l + TE < E ? e0l + TE : E);
  const int n = e1 > e0 ? (int)(e1 - e0) : 0;

  // round 1: slots (the slot array is padded to whole tiles), run bitmap and
  // bases, row bounds
  const uint32_t rb = tile_rb[t];
  const uint32_t re = (t + 1 == ntiles) ? V : tile_rb[t + 1];
  uint32_t sl[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) sl[j] = slot[e0l + j * kPullBlock + tid];
  for (int w = tid; w < W; w += kPullBlock) sW[w] = words[t * (uint64_t)W + w];
  for (int k = tid; k < K; k += kPullBlock) sD[k] = base[t * (uint64_t)K + k];
  if (tid == 0) nlong = 0;
  __syncthreads();

  // round 2: the staged sums, then the rows' offsets and data
  T sv[IPT];
  if (n > 0) {
    // run rank of stage entry i: set bits at or before i, minus one (entry 0
    // always starts a run; lanes past n take entry 0)
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const int i = j * kPullBlock + tid;
      const uint32_t ii = i < n ? (uint32_t)i : 0u;
      const uint2 wd = sW[ii >> 5];
      const uint32_t r = wd.y + (uint32_t)__popc(wd.x & ((2u << (ii & 31)) - 1u)) - 1u;
      sv[j] = m.gather(0, sD[r] + ii);
    }
  }
  TileRows<Mode, 4, Off> tr;
  tr.load(off, rb, re, m, tid);
  if (n > 0) {
#pragma unroll
    for (int j = 0; j < IPT; ++j) stage[j * kPullBlock + tid] = sv[j];
  }
  __syncthreads();
  if (n > 0) {
    T gv[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) gv[j] = stage[sl[j] & (uint32_t)(TE - 1)];
#pragma unroll
    for (int j = 0; j < IPT; ++j) vals[j * kPullBlock + tid] = gv[j];
  }
  __syncthreads();
  tile_rows<IPT>(t, e0, e1, n, off, rb, re, tr, carryP, carryO, part, m, vals, longj, &nlong, red, tid);
}

// Fix-up: tile t owns the row that starts in it and runs past its end.
template <int IPT, class Mode, class Off>
__global__ __launch_bounds__(kPullBlock) void k_pull_fixup(
    const Off* __restrict__ off, const uint32_t* __restrict__ tile_rb, uint64_t E, uint32_t V,
    uint64_t ntiles, const typename Mode::T* __restrict__ carryP,
    const typename Mode::T* __restrict__ carryO, double* __restrict__ part2, Mode m) {
  using T = typename Mode::T;
  constexpr uint64_t TE = kPullBlock * IPT;
  __shared__ double red[8];
  const uint64_t t = blockIdx.x * (uint64_t)kPullBlock + threadIdx.x;
  double acc0 = 0.0, acc1 = 0.0;
  if (t + 1 < ntiles) {
    const uint32_t rb = tile_rb[t], re = tile_rb[t + 1];
    const uint64_t e1 = (t + 1) * TE;
    if (re > rb && off[re] > e1) {
      const uint32_t r = re - 1;
      const uint64_t end = off[r + 1];
      T s = carryO[t];
      for (uint64_t q = t + 1; q * TE < end; ++q) s += carryP[q];
      m.finalize(r, s, acc0, acc1);
    }
  }
  block_sum2(acc0, acc1, red);
  if (threadIdx.x == 0) { part2[2 * blockIdx.x] = acc0; part2[2 * blockIdx.x + 1] = acc1; }
}

// Fixed-order reduction of n pairs into out[0], out[1] (single block).
__global__ void k_reduce_pairs(const double* __restrict__ a, uint64_t na, const double* __restrict__ b,
                               uint64_t nb, double* __restrict__ out);
// First level of the two-level form: block i writes its pair to out[2i..2i+1].
constexpr int kRedBlocks = 256;
__global__ void k_reduce_pairs_part(const double* __restrict__ a, uint64_t na, const double* __restrict__ b,
                                    uint64_t nb, double* __restrict__ out);
// Deterministic reduction of the pairs a[0..na) and b[0..nb) into out[0..1];
// tmp holds 2*kRedBlocks doubles.
void reduce_pairs(const double* a, uint64_t na, const double* b, uint64_t nb, double* out, double* tmp,
                  hipStream_t s);

// Host driver: one pull pass (tile + fix-up + reduce) over (off, idx).
struct PullPlan {
  const uint64_t* off = nullptr;
  const uint32_t* off32 = nullptr;  // used instead of off when set (every offset < 2^32)
  const uint32_t* idx = nullptr;
  const uint32_t* tile_rb = nullptr;
  uint64_t E = 0, ntiles = 0;
  uint32_t V = 0;
  int ipt = 8;
  bool nt = false;        // non-temporal hint on the neighbour-id stream
  void* carryP = nullptr;
  void* carryO = nullptr;
  double* part = nullptr;   // 2*ntiles
  double* part2 = nullptr;  // 2*ceil(ntiles/256)
  double* part3 = nullptr;  // 2*kRedBlocks
  const uint32_t* xmap = nullptr;  // class tile starts (source-range partitions) or nullptr
  // per-row pass over segment sums staged in LDS (k_pull_staged), or nullptr
  const uint16_t* slot = nullptr;
  const uint2* words = nullptr;
  const uint32_t* base = nullptr;
  int K = 0;
  unsigned grid = 0;               // workgroups of the tile kernel when xmap is set
  const char* tag = "pull_tile";   // profiling names of the two kernels
  const char* fix_tag = "pull_fixup";
};

// Fix-up of the rows that span tiles, then (scal_out != nullptr) the fixed-order
// reduction of the tiles' (acc0, acc1) pairs into scal_out.
template <int IPT, class Mode>
void pull_finish(const PullPlan& p, const Mode& m, double* scal_out, hipStream_t s) {
  using T = typename Mode::T;
  const unsigned nfb = (unsigned)((p.ntiles + kPullBlock - 1) / kPullBlock);
  {
    ProfScope ps(p.fix_tag);
    if (p.off32)
      hipLaunchKernelGGL((k_pull_fixup<IPT, Mode, uint32_t>), dim3(nfb), dim3(kPullBlock), 0, s, p.off32, p.tile_rb,
                         p.E, p.V, p.ntiles, (const T*)p.carryP, (const T*)p.carryO, p.part2, m);
    else
      hipLaunchKernelGGL((k_pull_fixup<IPT, Mode, uint64_t>), dim3(nfb), dim3(kPullBlock), 0, s, p.off, p.tile_rb,
                         p.E, p.V, p.ntiles, (const T*)p.carryP, (const T*)p.carryO, p.part2, m);
  }
  // scal_out == nullptr: the mode's finalize accumulates nothing (segment sums)
  if (scal_out) reduce_pairs(p.part, p.ntiles, p.part2, (uint64_t)nfb, scal_out, p.part3, s);
}

template <int IPT, class Mode, class Off>
void pull_launch(const PullPlan& p, const Off* off, const Mode& m, hipStream_t s) {
  using T = typename Mode::T;
  const unsigned grid = p.xmap ? p.grid : (unsigned)p.ntiles;
  if (p.slot) {
    if constexpr (IPT == 8)
      hipLaunchKernelGGL((k_pull_staged<IPT, Mode, Off>), dim3((unsigned)p.ntiles), dim3(kPullBlock), 0, s, off,
                         p.slot, p.tile_rb, p.E, p.V, p.ntiles, p.words, p.base, p.K, (T*)p.carryP, (T*)p.carryO,
                         p.part, m);
    else
      throw Error(GRE_DEV_NULL, "the staged per-row pass runs with 8 entries per thread");
  } else if (p.nt) {
    hipLaunchKernelGGL((k_pull_tile<IPT, true, Mode, Off>), dim3(grid), dim3(kPullBlock), 0, s, off, p.idx,
                       p.tile_rb, p.E, p.V, p.ntiles, (T*)p.carryP, (T*)p.carryO, p.part, p.xmap, m);
  } else {
    hipLaunchKernelGGL((k_pull_tile<IPT, false, Mode, Off>), dim3(grid), dim3(kPullBlock), 0, s, off, p.idx,
                       p.tile_rb, p.E, p.V, p.ntiles, (T*)p.carryP, (T*)p.carryO, p.part, p.xmap, m);
  }
}

template <int IPT, class Mode>
void pull_pass_ipt(const PullPlan& p, const Mode& m, double* scal_out, hipStream_t s) {
  {
    ProfScope ps(p.tag);
    if (p.off32) pull_launch<IPT>(p, p.off32, m, s);
    else pull_launch<IPT>(p, p.off, m, s);
  }
  pull_finish<IPT>(p, m, scal_out, s);
}

template <class Mode>
void pull_pass(const PullPlan& p, const Mode& m, double* scal_out, hipStream_t s) {
  switch (p.ipt) {
    case 4: pull_pass_ipt<4>(p, m, scal_out, s); break;
    case 8: pull_pass_ipt<8>(p, m, scal_out, s); break;
    case 16: pull_pass_ipt<16>(p, m, scal_out, s); break;
    default: throw Error(GRE_DEV_NULL, "tile items per thread must be 4, 8 or 16");
  }
  GRE_HIP(hipGetLastError());
}

// Allocates the plan's carry/partial buffers for a (graph, direction, ipt, T).
struct PullBuffers {
  DevBuf carryP, carryO, part, part2, part3;
  void make(PullPlan& p, gre_graph* g, bool csc, int ipt, size_t tbytes);
  // the same over the segments of a source-range partition
  void make(PullPlan& p, XPart& xp, int ipt, size_t tbytes);
  // and over its per-row segment lists (rows = the nV vertices)
  void make_pairs(PullPlan& p, XPart& xp, uint32_t nV, int ipt, size_t tbytes);
  void alloc_for(PullPlan& p, size_t tbytes);
};

}  // namespace gre
