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

template <int IPT, class Mode>
__device__ __forceinline__ void pull_tile_rows(
    uint64_t t, bool valid, uint64_t e0, uint64_t e1, int n, const uint64_t* __restrict__ off,
    const uint32_t* __restrict__ tile_rb, uint32_t V, uint64_t ntiles, typename Mode::T* __restrict__ carryP,
    typename Mode::T* __restrict__ carryO, double* __restrict__ part, const Mode& m, const typename Mode::T* vals,
    uint32_t* longj, int* nlong, double* red, int tid);

// One tile of TE = 256*IPT entries, processed by one 256-thread group (tid in
// [0, 256)).  vals/longj/nlong/red are the group's LDS.  gat(e, u) gives the
// value gathered for entry e with neighbour u.  valid = false: the group has no
// tile this round but still meets every block barrier.
template <int IPT, bool NT, class Mode, class Gather, class Id>
__device__ __forceinline__ void pull_tile_body(
    uint64_t t, bool valid, const uint64_t* __restrict__ off, const Id* __restrict__ idx,
    const uint32_t* __restrict__ tile_rb, uint64_t E, uint32_t V, uint64_t ntiles,
    typename Mode::T* __restrict__ carryP, typename Mode::T* __restrict__ carryO, double* __restrict__ part,
    const Mode& m, const Gather& gat, typename Mode::T* vals, uint32_t* longj, int* nlong, double* red,
    int tid) {
  using T = typename Mode::T;
  constexpr int TE = kPullBlock * IPT;
  const uint64_t e0 = t * (uint64_t)TE;
  const uint64_t e1 = valid ? (e0 + TE < E ? e0 + TE : E) : e0;
  const int n = e1 > e0 ? (int)(e1 - e0) : 0;

  // stage 1: neighbour ids (coalesced; NT = streamed with a non-temporal hint
  // so they do not push the gathered values out of the caches), then gathers
  uint32_t nb[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int i = j * kPullBlock + tid;
    if (NT) nb[j] = i < n ? (uint32_t)__builtin_nontemporal_load(idx + e0 + i) : 0u;
    else nb[j] = i < n ? (uint32_t)idx[e0 + i] : 0u;
  }
  // every gather is issued before any is used: a lane past the tile's end
  // reads entry e0 with neighbour 0 (nb = 0 there), a valid element, and
  // stores it to a slot of vals that no row reads (i < TE always).  Guarded
  // gathers or stores would make each load wait for the one before it.
  if (n > 0) {
    typename Mode::T gv[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const int i = j * kPullBlock + tid;
      gv[j] = gat(e0 + (i < n ? (uint64_t)i : 0ull), nb[j]);
    }
#pragma unroll
    for (int j = 0; j < IPT; ++j) vals[j * kPullBlock + tid] = gv[j];
  }
  if (tid == 0) *nlong = 0;
  __syncthreads();
  pull_tile_rows<IPT>(t, valid, e0, e1, n, off, tile_rb, V, ntiles, carryP, carryO, part, m, vals, longj, nlong, red,
                      tid);
}

// Stages 2 and 3 of a tile whose values vals[0..n) (entries e0..e1) are in
// LDS and *nlong = 0, both visible to the whole group.
template <int IPT, class Mode>
__device__ __forceinline__ void pull_tile_rows(
    uint64_t t, bool valid, uint64_t e0, uint64_t e1, int n, const uint64_t* __restrict__ off,
    const uint32_t* __restrict__ tile_rb, uint32_t V, uint64_t ntiles, typename Mode::T* __restrict__ carryP,
    typename Mode::T* __restrict__ carryO, double* __restrict__ part, const Mode& m, const typename Mode::T* vals,
    uint32_t* longj, int* nlong, double* red, int tid) {
  using T = typename Mode::T;
  // stage 2: rows that start in the tile, plus the row continuing into it
  double acc0 = 0.0, acc1 = 0.0;
  uint32_t rb = 0, nr = 0;
  bool has_rc = false;
  if (valid) {
    rb = tile_rb[t];
    const uint32_t re = (t + 1 == ntiles) ? V : tile_rb[t + 1];
    has_rc = rb > 0 && n > 0 && off[rb] > e0;
    nr = (re - rb) + (has_rc ? 1u : 0u);
  }
  for (uint32_t j = tid; j < nr; j += kPullBlock) {
    const uint32_t r = has_rc ? (j == 0 ? rb - 1 : rb + j - 1) : rb + j;
    const uint64_t ob = off[r], oe = off[r + 1];
    const uint64_t sb = ob > e0 ? ob : e0;
    const uint64_t se = oe < e1 ? oe : e1;
    const int len = se > sb ? (int)(se - sb) : 0;
    if (len > kLongRow) {
      const int k = atomicAdd(nlong, 1);
      longj[k] = j;
      continue;
    }
    T s = (T)0;
    const int base = (int)(sb - e0);
    for (int q = 0; q < len; ++q) s += vals[base + q];
    if (ob < e0) carryP[t] = s;
    else if (oe > e1) carryO[t] = s;
    else m.finalize(r, s, acc0, acc1);
  }
  __syncthreads();

  // stage 3: long rows, one wave each
  const int w = tid >> 6, l = tid & 63;
  const int nl = *nlong;
  for (int q = w; q < nl; q += kPullBlock / 64) {
    const uint32_t j = longj[q];
    const uint32_t r = has_rc ? (j == 0 ? rb - 1 : rb + j - 1) : rb + j;
    const uint64_t ob = off[r], oe = off[r + 1];
    const uint64_t sb = ob > e0 ? ob : e0;
    const uint64_t se = oe < e1 ? oe : e1;
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
  if (valid && tid == 0) { part[2 * t] = acc0; part[2 * t + 1] = acc1; }
}

template <class Mode>
struct ModeGather {
  const Mode& m;
  __device__ typename Mode::T operator()(uint64_t e, uint32_t u) const { return m.gather(e, u); }
};

template <int IPT, bool NT, class Mode>
__global__ __launch_bounds__(kPullBlock) void k_pull_tile(
    const uint64_t* __restrict__ off, const uint32_t* __restrict__ idx,
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
  pull_tile_body<IPT, NT>(t, true, off, idx, tile_rb, E, V, ntiles, carryP, carryO, part, m, ModeGather<Mode>{m},
                          vals, longj, &nlong, red, (int)threadIdx.x);
}

// The per-row pass over a partition's segment sums (rows = vertices, entries =
// each row's segment ids).  A tile's segments form one run of consecutive ids
// per (class, range) key; the tile first copies those runs from the mode's
// gather array into LDS (coalesced: entry i of the stage belongs to the run k
// with runO[k] <= i < runO[k+1], at segment runA[k] + i - runO[k]), then pulls
// with each entry's value read from its LDS slot.  The entries, and so the
// order of every row's sum, are those of the plain tile pull over the segment
// ids.  The loads of the tile are issued in two rounds: (slots, run table, row
// bounds), then (staged sums, row offsets, each row's own data through the
// mode's pre(r)), so a tile waits for two memory latencies, not one per step.
// Thread j takes row rb - 1 + j: row rb - 1 only when it continues into the
// tile.  Modes provide pre(r) and finalize_pre(r, s, pre, acc0, acc1).
template <class M, class = void>
struct HasPre : std::false_type {};
template <class M>
struct HasPre<M, std::void_t<decltype(std::declval<const M&>().pre(0u))>> : std::true_type {};

constexpr int kStagedRows = 4;  // rows per thread whose data is loaded with the stage

template <int IPT, class Mode>
__global__ __launch_bounds__(kPullBlock) void k_pull_staged(
    const uint64_t* __restrict__ off, const uint16_t* __restrict__ slot, const uint32_t* __restrict__ tile_rb,
    uint64_t E, uint32_t V, uint64_t ntiles, const uint32_t* __restrict__ runA, const uint16_t* __restrict__ runO,
    int K, typename Mode::T* __restrict__ carryP, typename Mode::T* __restrict__ carryO, double* __restrict__ part,
    Mode m) {
  using T = typename Mode::T;
  using Pre = decltype(m.pre(0u));
  constexpr int TE = kPullBlock * IPT;
  constexpr int Q = kStagedRows;
  static_assert((TE & (TE - 1)) == 0, "tile entries must be a power of two");
  __shared__ T stage[TE];
  __shared__ T vals[TE];
  __shared__ uint32_t sA[kMaxKeys];
  __shared__ uint32_t sO[2 * kMaxKeys + 1];
  __shared__ uint32_t longj[TE / kLongRow + 2];
  __shared__ int nlong;
  __shared__ double red[8];
  const uint64_t t = blockIdx.x;
  const int tid = (int)threadIdx.x;
  const uint64_t e0 = t * (uint64_t)TE;
  const uint64_t e1 = e0 + TE < E ? e0 + TE : E;
  const int n = e1 > e0 ? (int)(e1 - e0) : 0;

  // round 1: slots (the slot array is padded to whole tiles), run table, row bounds
  const uint32_t rb = tile_rb[t];
  const uint32_t re = (t + 1 == ntiles) ? V : tile_rb[t + 1];
  uint32_t sl[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) sl[j] = slot[e0 + j * kPullBlock + tid];
  // sO[k] for k in (K, 2K] is a sentinel above every entry index, so the
  // search below needs no bound test
  for (int k = tid; k <= 2 * K; k += kPullBlock) {
    sO[k] = k <= K ? (uint32_t)runO[t * (uint64_t)(K + 1) + k] : 0xffffffffu;
    if (k < K) sA[k] = runA[t * (uint64_t)K + k];
  }
  if (tid == 0) nlong = 0;
  __syncthreads();

  // round 2: staged sums, then each of the thread's first Q rows' offsets and data
  T sv[IPT];
  if (n > 0) {
    // run of stage entry i: the largest k with sO[k] <= i, for all IPT entries
    // at once (lanes past n take entry 0)
    uint32_t ii[IPT];
    uint32_t lo[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const int i = j * kPullBlock + tid;
      ii[j] = i < n ? (uint32_t)i : 0u;
      lo[j] = 0;
    }
    int b = 1;
    while (2 * b < K) b *= 2;
    for (; b > 0; b >>= 1) {
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t c = lo[j] + (uint32_t)b;
        lo[j] = sO[c] <= ii[j] ? c : lo[j];
      }
    }
#pragma unroll
    for (int j = 0; j < IPT; ++j) sv[j] = m.gather(0, sA[lo[j]] + (ii[j] - sO[lo[j]]));
  }
  uint64_t rob[Q], roe[Q];
  Pre rpre[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int64_t r = (int64_t)rb - 1 + tid + q * kPullBlock;
    const uint32_t rs = (r >= 0 && r < (int64_t)re) ? (uint32_t)r : 0u;
    rob[q] = off[rs];
    roe[q] = off[rs + 1];
    rpre[q] = m.pre(rs);
  }
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

  // rows rb - 1 .. re - 1; a row longer than kLongRow in-tile entries goes to a wave
  const uint32_t nrows = re - rb + 1;
  // row rb - 1 is thread 0's first row: roe[0] = off[rb] there
  const bool has_rc = rb > 0 && n > 0 && roe[0] > e0;
  double acc0 = 0.0, acc1 = 0.0;
  auto row = [&](uint32_t j, uint64_t ob, uint64_t oe, const Pre& pr) {
    const uint32_t r = rb - 1 + j;
    const uint64_t sb = ob > e0 ? ob : e0;
    const uint64_t se = oe < e1 ? oe : e1;
    const int len = se > sb ? (int)(se - sb) : 0;
    if (len > kLongRow) {
      longj[atomicAdd(&nlong, 1)] = j;
      return;
    }
    T s = (T)0;
    const int base = (int)(sb - e0);
    for (int q = 0; q < len; ++q) s += vals[base + q];
    if (ob < e0) carryP[t] = s;
    else if (oe > e1) carryO[t] = s;
    else m.finalize_pre(r, s, pr, acc0, acc1);
  };
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const uint32_t j = (uint32_t)(tid + q * kPullBlock);
    if (j < nrows && (j > 0 || has_rc)) row(j, rob[q], roe[q], rpre[q]);
  }
  for (uint32_t j = (uint32_t)(tid + Q * kPullBlock); j < nrows; j += kPullBlock) {
    const uint32_t r = rb - 1 + j;
    row(j, off[r], off[r + 1], m.pre(r));
  }
  __syncthreads();

  // long rows, one wave each
  const int w = tid >> 6, l = tid & 63;
  const int nl = nlong;
  for (int q = w; q < nl; q += kPullBlock / 64) {
    const uint32_t r = rb - 1 + longj[q];
    const uint64_t ob = off[r], oe = off[r + 1];
    const uint64_t sb = ob > e0 ? ob : e0;
    const uint64_t se = oe < e1 ? oe : e1;
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

// Fix-up: tile t owns the row that starts in it and runs past its end.
template <int IPT, class Mode>
__global__ __launch_bounds__(kPullBlock) void k_pull_fixup(
    const uint64_t* __restrict__ off, const uint32_t* __restrict__ tile_rb, uint64_t E, uint32_t V,
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
  const uint32_t* runA = nullptr;
  const uint16_t* runO = nullptr;
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
    hipLaunchKernelGGL((k_pull_fixup<IPT, Mode>), dim3(nfb), dim3(kPullBlock), 0, s, p.off, p.tile_rb, p.E, p.V,
                       p.ntiles, (const T*)p.carryP, (const T*)p.carryO, p.part2, m);
  }
  // scal_out == nullptr: the mode's finalize accumulates nothing (segment sums)
  if (scal_out) reduce_pairs(p.part, p.ntiles, p.part2, (uint64_t)nfb, scal_out, p.part3, s);
}

template <int IPT, class Mode>
void pull_pass_ipt(const PullPlan& p, const Mode& m, double* scal_out, hipStream_t s) {
  using T = typename Mode::T;
  {
    ProfScope ps(p.tag);
    const unsigned grid = p.xmap ? p.grid : (unsigned)p.ntiles;
    if (p.slot) {
      if constexpr (HasPre<Mode>::value)
        hipLaunchKernelGGL((k_pull_staged<IPT, Mode>), dim3((unsigned)p.ntiles), dim3(kPullBlock), 0, s, p.off,
                           p.slot, p.tile_rb, p.E, p.V, p.ntiles, p.runA, p.runO, p.K, (T*)p.carryP, (T*)p.carryO,
                           p.part, m);
      else
        throw Error(GRE_DEV_NULL, "this pull has no staged per-row form");
    } else if (p.nt)
      hipLaunchKernelGGL((k_pull_tile<IPT, true, Mode>), dim3(grid), dim3(kPullBlock), 0, s, p.off, p.idx,
                         p.tile_rb, p.E, p.V, p.ntiles, (T*)p.carryP, (T*)p.carryO, p.part, p.xmap, m);
    else
      hipLaunchKernelGGL((k_pull_tile<IPT, false, Mode>), dim3(grid), dim3(kPullBlock), 0, s, p.off, p.idx,
                         p.tile_rb, p.E, p.V, p.ntiles, (T*)p.carryP, (T*)p.carryO, p.part, p.xmap, m);
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
