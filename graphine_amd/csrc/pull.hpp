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
template <int IPT, bool NT, class Mode>
__global__ __launch_bounds__(kPullBlock) void k_pull_tile(
    const uint64_t* __restrict__ off, const uint32_t* __restrict__ idx,
    const uint32_t* __restrict__ tile_rb, uint64_t E, uint32_t V, uint64_t ntiles,
    typename Mode::T* __restrict__ carryP, typename Mode::T* __restrict__ carryO,
    double* __restrict__ part, Mode m) {
  using T = typename Mode::T;
  constexpr int TE = kPullBlock * IPT;
  __shared__ T vals[TE];
  __shared__ uint32_t longj[TE / kLongRow + 2];
  __shared__ int nlong;
  __shared__ double red[8];

  const uint64_t t = blockIdx.x;
  const uint64_t e0 = t * (uint64_t)TE;
  const uint64_t e1 = e0 + TE < E ? e0 + TE : E;
  const int n = e1 > e0 ? (int)(e1 - e0) : 0;
  const int tid = threadIdx.x;

  // stage 1: neighbour ids (coalesced; NT = streamed with a non-temporal hint
  // so they do not push the gathered values out of the caches), then gathers
  uint32_t nb[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int i = j * kPullBlock + tid;
    if (NT) nb[j] = i < n ? __builtin_nontemporal_load(idx + e0 + i) : 0u;
    else nb[j] = i < n ? idx[e0 + i] : 0u;
  }
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const int i = j * kPullBlock + tid;
    if (i < n) vals[i] = m.gather(e0 + i, nb[j]);
  }
  if (tid == 0) nlong = 0;
  __syncthreads();

  // stage 2: rows that start in the tile, plus the row continuing into it
  const uint32_t rb = tile_rb[t];
  const uint32_t re = (t + 1 == ntiles) ? V : tile_rb[t + 1];
  const bool has_rc = rb > 0 && n > 0 && off[rb] > e0;
  const uint32_t nr = (re - rb) + (has_rc ? 1u : 0u);
  double acc0 = 0.0, acc1 = 0.0;
  for (uint32_t j = tid; j < nr; j += kPullBlock) {
    const uint32_t r = has_rc ? (j == 0 ? rb - 1 : rb + j - 1) : rb + j;
    const uint64_t ob = off[r], oe = off[r + 1];
    const uint64_t sb = ob > e0 ? ob : e0;
    const uint64_t se = oe < e1 ? oe : e1;
    const int len = se > sb ? (int)(se - sb) : 0;
    if (len > kLongRow) {
      const int k = atomicAdd(&nlong, 1);
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
  const int nl = nlong;
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

  block_sum2(acc0, acc1, red);
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
};

template <int IPT, class Mode>
void pull_pass_ipt(const PullPlan& p, const Mode& m, double* scal_out, hipStream_t s) {
  using T = typename Mode::T;
  {
    ProfScope ps("pull_tile");
    if (p.nt)
      hipLaunchKernelGGL((k_pull_tile<IPT, true, Mode>), dim3((unsigned)p.ntiles), dim3(kPullBlock), 0, s, p.off,
                         p.idx, p.tile_rb, p.E, p.V, p.ntiles, (T*)p.carryP, (T*)p.carryO, p.part, m);
    else
      hipLaunchKernelGGL((k_pull_tile<IPT, false, Mode>), dim3((unsigned)p.ntiles), dim3(kPullBlock), 0, s, p.off,
                         p.idx, p.tile_rb, p.E, p.V, p.ntiles, (T*)p.carryP, (T*)p.carryO, p.part, m);
  }
  const unsigned nfb = (unsigned)((p.ntiles + kPullBlock - 1) / kPullBlock);
  {
    ProfScope ps("pull_fixup");
    hipLaunchKernelGGL((k_pull_fixup<IPT, Mode>), dim3(nfb), dim3(kPullBlock), 0, s, p.off, p.tile_rb, p.E, p.V,
                       p.ntiles, (const T*)p.carryP, (const T*)p.carryO, p.part2, m);
  }
  reduce_pairs(p.part, p.ntiles, p.part2, (uint64_t)nfb, scal_out, p.part3, s);
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
};

}  // namespace gre
