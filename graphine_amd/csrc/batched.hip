// Batched personalised PageRank (k vectors in one launch sequence), its
// adjoint and the batched edge gradient.
//
// Vectors are stored vertex-major (nV x k), so one neighbour gather fetches
// k contiguous values.  A pass is three kernels:
//   walk   a group of G = pow2ceil(k) lanes walks one tile of kTileG adjacency
//          entries: lanes load G neighbour ids at a time (coalesced), issue G
//          independent k-wide gathers, then fold them into a per-lane running
//          sum and store each row's raw sum when the row ends.  Row ends come
//          from a window of G offsets held one per lane (one coalesced load
//          per G rows), so nothing in the fold waits on memory;
//   fix-up rows crossing a tile boundary add their partial vectors in tile
//          order, so results are deterministic;
//   apply  one coalesced pass over the nV x k sums applies the update and
//          forms the per-vector FP64 reductions (dangling mass, p.lambda, L1
//          residual) as fixed-order block partials.
#include "pr.hpp"

#include <algorithm>
#include <cstring>
#include <initializer_list>

namespace gre {
namespace {

constexpr int kTileG = 1024;   // adjacency entries per group tile
constexpr int kBlk = 256;

unsigned gcap(uint64_t n) {
  uint64_t g = (n + kBlk - 1) / kBlk;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

int pow2ceil(int k) {
  int g = 1;
  while (g < k) g <<= 1;
  return g;
}

template <typename TT>
struct BFwd {  // forward: gather x (CSC), finalise ranks (raw sums go to xout)
  const TT* xin; TT* xout; const TT* w; const TT* p; const int32_t* outdeg;
  const double* scal;  // [2k]: dang_j at scal[2j]
  double damp; int k;
  bool resid = true;   // accumulate the L1 change (fixed-count runs need it on the last iteration only)
  __device__ TT* sums() const { return xout; }
  __device__ TT gather(uint64_t e, uint32_t u, int j) const { return (TT)w[e] * xin[(uint64_t)u * k + j]; }
  __device__ TT wt(uint64_t e) const { return (TT)w[e]; }
  __device__ TT val(uint32_t u, int j) const { return xin[(uint64_t)u * k + j]; }
  __device__ void finalize(uint32_t r, int j, TT s, double& a0, double& a1) const {
    const uint64_t o = (uint64_t)r * k + j;
    const TT pv = p[o];
    const TT y = (TT)damp * (s + (TT)scal[2 * j] * pv) + (TT)(1.0 - damp) * pv;
    if (resid) a1 += fabs((double)y - (double)xin[o]);
    xout[o] = y;
    if (outdeg[r] == 0) a0 += (double)y;
  }
};

template <typename TT>
struct BAdj {  // adjoint: gather lambda (CSR), finalise lambda (raw sums go to wl)
  const TT* lin; TT* lout; const TT* w; const TT* c; const TT* p; const int32_t* outdeg; TT* wl;
  const double* scal;  // [2k]: p.lambda_j at scal[2j]
  double damp; int k;
  bool resid = true;   // accumulate the L1 change (fixed-count runs need it on the last iteration only)
  __device__ TT* sums() const { return wl; }
  __device__ TT gather(uint64_t e, uint32_t v, int j) const { return (TT)w[e] * lin[(uint64_t)v * k + j]; }
  __device__ TT wt(uint64_t e) const { return (TT)w[e]; }
  __device__ TT val(uint32_t v, int j) const { return lin[(uint64_t)v * k + j]; }
  __device__ void finalize(uint32_t u, int j, TT s, double& a0, double& a1) const {
    // s was read from wl (sums()), where the pass left the raw sum for the gradients
    const uint64_t o = (uint64_t)u * k + j;
    const TT y = c[o] + (TT)damp * (s + (outdeg[u] == 0 ? (TT)scal[2 * j] : (TT)0));
    if (resid) a1 += fabs((double)y - (double)lin[o]);
    lout[o] = y;
    a0 += (double)p[o] * (double)y;
  }
};

// Group-walk tile kernel.  Group size G (lanes), groups per block kBlk/G.
template <int G, class M, typename TT>
__global__ __launch_bounds__(kBlk) void k_bwalk(const uint64_t* __restrict__ off, const uint32_t* __restrict__ idx,
                                                const uint32_t* __restrict__ tile_rb, uint64_t E, uint32_t V,
                                                uint64_t ntiles, TT* __restrict__ carryP, TT* __restrict__ carryO,
                                                M m) {
  constexpr int NG = kBlk / G;
  const int gl = threadIdx.x % G, gi = threadIdx.x / G;
  const int k = m.k;
  const bool act = gl < k;
  const uint64_t t = (uint64_t)blockIdx.x * NG + gi;
  if (t >= ntiles || V == 0) return;
  TT* __restrict__ sums = m.sums();
  const uint64_t e0 = t * (uint64_t)kTileG;
  const uint64_t e1 = e0 + kTileG < E ? e0 + kTileG : E;
  const uint32_t rb = tile_rb[t];
  const uint32_t re = (t + 1 == ntiles) ? V : tile_rb[t + 1];
  uint32_t r = (rb > 0 && e1 > e0 && off[rb] > e0) ? rb - 1 : rb;
  // row ends of rows wr .. wr+G-1, one per lane (E past the last row)
  uint32_t wr = r;
  uint64_t endw = (uint64_t)wr + 1 + gl <= V ? off[(uint64_t)wr + 1 + gl] : E;
  uint64_t rbeg = r < V ? off[r] : E;
  uint64_t rend = __shfl(endw, 0, G);
  TT acc = (TT)0;
  // stores row r's sum (or a carry) and moves to the next row
  auto flush_next = [&]() {
    if (act) {
      if (rbeg < e0) carryP[t * k + gl] = acc;
      else if (rend > e1) carryO[t * k + gl] = acc;
      else sums[(uint64_t)r * k + gl] = acc;
    }
    acc = (TT)0;
    rbeg = rend;
    ++r;
    if (r - wr >= (uint32_t)G) {
      wr = r;
      endw = (uint64_t)wr + 1 + gl <= V ? off[(uint64_t)wr + 1 + gl] : E;
    }
    rend = __shfl(endw, (int)(r - wr), G);
  };
  // ids and weights of the chunk after the current one are loaded before the
  // current chunk is folded
  uint32_t myn = e0 + gl < e1 ? idx[e0 + gl] : 0u;
  TT mywn = e0 + gl < e1 ? m.wt(e0 + gl) : (TT)0;
  for (uint64_t cb = e0; cb < e1; cb += G) {
    // the group's next G entries: ids and weights one per lane (coalesced),
    // then G gathers of k contiguous values; each entry's id and weight reach
    // the group by shuffles, so a lane keeps one address per gather in flight
    const uint32_t myu = myn;
    const TT myw = mywn;
    // unconditional gathers (past the tile's end: neighbour 0; idle lanes:
    // component 0), so all G are in flight before the first use
    TT g[G];
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const uint32_t u = __shfl(myu, q, G);
      const bool ok = cb + q < e1;
      const TT v = m.val(u, act ? gl : 0);
      g[q] = (act && ok) ? v : (TT)0;
    }
    {
      const uint64_t nx = cb + G + gl;
      myn = nx < e1 ? idx[nx] : 0u;
      mywn = nx < e1 ? m.wt(nx) : (TT)0;
    }
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const uint64_t e = cb + q;
      if (e >= e1) break;
      while (e >= rend) flush_next();  // rows ending before e, empty rows included
      acc += __shfl(myw, q, G) * g[q];
    }
  }
  // rows still open at the tile end, plus empty rows that start in the tile
  while (r < re) flush_next();
}

// Group-per-row kernel (GRE_SCHED_ROWS): a group of G lanes takes whole
// destination rows (grid-stride), G entries at a time with the next chunk's
// ids prefetched; no tile splits, so no carries.  Rows are summed in row order.
template <int G, class M, typename TT>
__global__ __launch_bounds__(kBlk) void k_brows(const uint64_t* __restrict__ off, const uint32_t* __restrict__ idx,
                                                uint32_t V, M m) {
  constexpr int NG = kBlk / G;
  const int gl = threadIdx.x % G, gi = threadIdx.x / G;
  const int k = m.k;
  const bool act = gl < k;
  TT* __restrict__ sums = m.sums();
  for (uint64_t r = (uint64_t)blockIdx.x * NG + gi; r < V; r += (uint64_t)gridDim.x * NG) {
    const uint64_t b = off[r], e = off[r + 1];
    TT acc = (TT)0;
    uint32_t myn = b + gl < e ? idx[b + gl] : 0u;
    TT mywn = b + gl < e ? m.wt(b + gl) : (TT)0;
    for (uint64_t cb = b; cb < e; cb += G) {
      const uint32_t myu = myn;
      const TT myw = mywn;
      TT g[G];
#pragma unroll
      for (int q = 0; q < G; ++q) {
        const uint32_t u = __shfl(myu, q, G);
        const bool ok = cb + q < e;
        const TT v = m.val(u, act ? gl : 0);
        g[q] = (act && ok) ? v : (TT)0;
      }
      {
        const uint64_t nx = cb + G + gl;
        myn = nx < e ? idx[nx] : 0u;
        mywn = nx < e ? m.wt(nx) : (TT)0;
      }
#pragma unroll
      for (int q = 0; q < G; ++q) acc += __shfl(myw, q, G) * g[q];
    }
    if (act) sums[r * k + gl] = acc;
  }
}

// rows that start in tile t and run past its end: carryO[t] + carryP of the
// tiles they continue into, in tile order
template <int G, class M, typename TT>
__global__ __launch_bounds__(kBlk) void k_bfix(const uint64_t* __restrict__ off, const uint32_t* __restrict__ tile_rb,
                                               uint64_t E, uint32_t V, uint64_t ntiles,
                                               const TT* __restrict__ carryP, const TT* __restrict__ carryO, M m) {
  constexpr int NG = kBlk / G;
  const int gl = threadIdx.x % G, gi = threadIdx.x / G;
  const int k = m.k;
  const uint64_t t = (uint64_t)blockIdx.x * NG + gi;
  if (t + 1 < ntiles && gl < k) {
    const uint32_t rb = tile_rb[t], re = tile_rb[t + 1];
    const uint64_t e1 = (t + 1) * (uint64_t)kTileG;
    if (re > rb && off[re] > e1) {
      const uint32_t r = re - 1;
      const uint64_t end = off[r + 1];
      TT s = carryO[t * k + gl];
      for (uint64_t q = t + 1; q * kTileG < end; ++q) s += carryP[q * k + gl];
      m.sums()[(uint64_t)r * k + gl] = s;
    }
  }
}

// apply: every row's update from its raw sum; per-vector block partials of
// (a0, a1) in part[2j][nblk], part[2j+1][nblk]
template <int G, class M, typename TT>
__global__ __launch_bounds__(kBlk) void k_bapply(uint32_t V, double* __restrict__ part, uint64_t nblk, M m) {
  constexpr int NG = kBlk / G;
  __shared__ double red0[NG][G], red1[NG][G];
  const int gl = threadIdx.x % G, gi = threadIdx.x / G;
  const int k = m.k;
  double a0 = 0.0, a1 = 0.0;
  if (gl < k) {
    const TT* __restrict__ sums = m.sums();
    for (uint64_t r = (uint64_t)blockIdx.x * NG + gi; r < V; r += (uint64_t)gridDim.x * NG)
      m.finalize((uint32_t)r, gl, sums[r * k + gl], a0, a1);
  }
  red0[gi][gl] = a0;
  red1[gi][gl] = a1;
  __syncthreads();
  if (threadIdx.x < (unsigned)k) {
    double s0 = 0.0, s1 = 0.0;
    for (int q = 0; q < NG; ++q) { s0 += red0[q][threadIdx.x]; s1 += red1[q][threadIdx.x]; }
    part[(uint64_t)(2 * threadIdx.x) * nblk + blockIdx.x] = s0;
    part[(uint64_t)(2 * threadIdx.x + 1) * nblk + blockIdx.x] = s1;
  }
}

// out[o] = sum of a[o*na + i] + b[o*nb + i] (one block per output, fixed order)
__global__ void k_reduce_rows(const double* __restrict__ a, uint64_t na, const double* __restrict__ b, uint64_t nb,
                              double* __restrict__ out) {
  __shared__ double sh[256];
  const int o = blockIdx.x;
  double s = 0.0;
  for (uint64_t i = threadIdx.x; i < na; i += blockDim.x) s += a[(uint64_t)o * na + i];
  for (uint64_t i = threadIdx.x; i < nb; i += blockDim.x) s += b[(uint64_t)o * nb + i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (unsigned w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[o] = sh[0];
}

// init: x = p (forward) or lambda = c (adjoint); a0_j = sum over dangling of x_j
// (forward) or p_j . c_j (adjoint).  part is [2k][nblk].
template <typename TT>
__global__ void k_binit(uint32_t nV, int k, const TT* __restrict__ src, const TT* __restrict__ pw,
                        const int32_t* __restrict__ outdeg, TT* __restrict__ dst, int adjoint,
                        double* __restrict__ part, uint64_t nblk) {
  __shared__ double sh[kBlk];
  // thread layout: lanes over j (k <= 64), rows strided
  const int G = 64;
  const int j = threadIdx.x % G, rl = threadIdx.x / G;
  double a0 = 0.0;
  if (j < k) {
    for (uint64_t v = (uint64_t)blockIdx.x * (kBlk / G) + rl; v < nV; v += (uint64_t)gridDim.x * (kBlk / G)) {
      const uint64_t o = v * k + j;
      const TT s = src[o];
      dst[o] = s;
      if (adjoint) a0 += (double)pw[o] * (double)s;
      else if (outdeg[v] == 0) a0 += (double)s;
    }
  }
  sh[threadIdx.x] = a0;
  __syncthreads();
  if (threadIdx.x < (unsigned)k) {
    double s = 0.0;
    for (int q = 0; q < kBlk / G; ++q) s += sh[q * G + threadIdx.x];
    part[(uint64_t)(2 * threadIdx.x) * nblk + blockIdx.x] = s;
    part[(uint64_t)(2 * threadIdx.x + 1) * nblk + blockIdx.x] = 0.0;
  }
}

template <typename TT>
__global__ void k_btargets(const uint32_t* __restrict__ tg, int64_t nt, int k, TT* __restrict__ c) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nt * k; i += (int64_t)gridDim.x * blockDim.x)
    c[(uint64_t)tg[i / k] * k + (i % k)] = (TT)1;
}

// batched edge gradient: one group per tile of CSR entries.  Per chunk of G
// entries the lambda rows of all G destinations are gathered before the first
// is used; rows without edges are skipped through a window of row ends held
// one per lane, and x, wl and the row sum are loaded once per row with edges.
template <int G, typename TT>
__global__ __launch_bounds__(kBlk) void k_bgrad(const uint64_t* __restrict__ off, const uint32_t* __restrict__ idx,
                                                const uint32_t* __restrict__ tile_rb, uint64_t E, uint32_t V,
                                                uint64_t ntiles, int k, const double* __restrict__ rowsum,
                                                const TT* __restrict__ x, const TT* __restrict__ lam,
                                                const TT* __restrict__ wl, double damp, float* __restrict__ gW,
                                                float* __restrict__ graw) {
  constexpr int NG = kBlk / G;
  const int gl = threadIdx.x % G, gi = threadIdx.x / G;
  const bool act = gl < k;
  const int jj = act ? gl : 0;
  const uint64_t t = (uint64_t)blockIdx.x * NG + gi;
  if (t >= ntiles) return;
  const uint64_t e0 = t * (uint64_t)kTileG;
  const uint64_t e1 = e0 + kTileG < E ? e0 + kTileG : E;
  if (e1 <= e0) return;
  const uint32_t rb = tile_rb[t];
  uint32_t r = (rb > 0 && off[rb] > e0) ? rb - 1 : rb;
  // row ends of rows wr .. wr+G-1, one per lane
  uint32_t wr = r;
  uint64_t endw = (uint64_t)wr + 1 + gl <= V ? off[(uint64_t)wr + 1 + gl] : E;
  uint64_t rend = __shfl(endw, 0, G);
  double xu = 0.0, xwl = 0.0, rs = 0.0;
  // moves r to the row holding entry e and loads that row's x, x.wl and row sum
  auto seek = [&](uint64_t e) {
    while (e >= rend) {
      ++r;
      if (r - wr >= (uint32_t)G) {
        wr = r;
        endw = (uint64_t)wr + 1 + gl <= V ? off[(uint64_t)wr + 1 + gl] : E;
      }
      rend = __shfl(endw, (int)(r - wr), G);
    }
    const uint64_t o = (uint64_t)r * k + jj;
    const double xv = (double)x[o], wv = (double)wl[o];
    rs = rowsum[r];
    xu = act ? xv : 0.0;
    double p = act ? xv * wv : 0.0;
    for (int sh = G / 2; sh > 0; sh >>= 1) p += __shfl_xor(p, sh, G);
    xwl = p;
  };
  seek(e0);
  for (uint64_t cb = e0; cb < e1; cb += G) {
    const uint64_t my = cb + gl;
    const uint32_t myv = my < e1 ? idx[my] : 0u;
    TT lv[G];
#pragma unroll
    for (int q = 0; q < G; ++q) lv[q] = lam[(uint64_t)__shfl(myv, q, G) * k + jj];
    float outW = 0.f, outR = 0.f;
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const uint64_t e = cb + q;
      if (e < e1) {
        if (e >= rend) seek(e);
        double d = act ? xu * (double)lv[q] : 0.0;
        for (int sh = G / 2; sh > 0; sh >>= 1) d += __shfl_xor(d, sh, G);
        if (gl == q) {
          outW = (float)(damp * d);
          outR = rs > 0.0 ? (float)(damp * (d - xwl) / rs) : 0.f;
        }
      }
    }
    if (my < e1) {
      gW[my] = outW;
      if (graw) graw[my] = outR;
    }
  }
}

template <typename TT>
__global__ void k_bgrad_p(uint32_t nV, int k, const TT* __restrict__ lam, const double* __restrict__ dang, double damp,
                          TT* __restrict__ gp) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < (uint64_t)nV * k;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const int j = (int)(i % k);
    gp[i] = (TT)(((1.0 - damp) + damp * dang[j]) * (double)lam[i]);
  }
}

size_t tsz(const gre_pr* P) { return P->dtype == GRE_F64 ? 8 : 4; }

struct BPlan {
  const uint64_t* off; const uint32_t* idx; const uint32_t* trb;
  uint64_t E, ntiles; uint32_t V;
  unsigned nblk, nfix, napply;   // walk and fix-up blocks (one tile per group), apply blocks
};

BPlan make_plan(gre_pr* P, bool csc) {
  gre_graph* g = P->g;
  BPlan b;
  b.off = (csc ? g->coff : g->off).as<uint64_t>();
  b.idx = (csc ? g->cidx : g->idx).as<uint32_t>();
  b.trb = graph_tiles(g, csc, kTileG, &b.ntiles);
  b.E = g->nE;
  b.V = g->nV;
  const int G = pow2ceil(P->k);
  const int NG = kBlk / G;
  b.nblk = (unsigned)((b.ntiles + NG - 1) / NG);
  b.nfix = b.nblk;
  const uint64_t nrowblk = ((uint64_t)b.V + NG - 1) / NG;
  b.napply = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(nrowblk, 16384));
  BatchState& bs = P->bstate;
  const size_t vec = tsz(P) * (size_t)P->k;
  if (bs.ntiles < b.ntiles) {
    bs.carryP.alloc(vec * b.ntiles);
    bs.carryO.alloc(vec * b.ntiles);
    bs.ntiles = b.ntiles;
  }
  const size_t need = sizeof(double) * 2 * (size_t)P->k * b.napply;
  if (bs.part.bytes < need) bs.part.alloc(need);
  return b;
}

void ensure_bscal(gre_pr* P, int iters) {
  BatchState& bs = P->bstate;
  if (bs.scal_cap >= iters + 2) return;
  bs.scal_cap = iters + 2;
  bs.scal.alloc(sizeof(double) * 2 * P->k * (size_t)bs.scal_cap);
}

template <int G, class M, typename TT>
void bpass(const BPlan& b, gre_pr* P, const M& m, double* scal_out, hipStream_t s, int sched) {
  BatchState& bs = P->bstate;
  if (sched == GRE_SCHED_ROWS) {
    ProfScope ps("batch_walk");
    hipLaunchKernelGGL((k_brows<G, M, TT>), dim3(b.napply), dim3(kBlk), 0, s, b.off, b.idx, b.V, m);
  } else {
    {
      ProfScope ps("batch_walk");
      hipLaunchKernelGGL((k_bwalk<G, M, TT>), dim3(b.nblk), dim3(kBlk), 0, s, b.off, b.idx, b.trb, b.E, b.V,
                         b.ntiles, bs.carryP.as<TT>(), bs.carryO.as<TT>(), m);
    }
    {
      ProfScope ps("batch_fixup");
      hipLaunchKernelGGL((k_bfix<G, M, TT>), dim3(b.nfix), dim3(kBlk), 0, s, b.off, b.trb, b.E, b.V, b.ntiles,
                         bs.carryP.as<TT>(), bs.carryO.as<TT>(), m);
    }
  }
  {
    ProfScope ps("batch_apply");
    hipLaunchKernelGGL((k_bapply<G, M, TT>), dim3(b.napply), dim3(kBlk), 0, s, b.V, bs.part.as<double>(),
                       (uint64_t)b.napply, m);
  }
  hipLaunchKernelGGL(k_reduce_rows, dim3(2 * P->k), dim3(256), 0, s, bs.part.as<double>(), (uint64_t)b.napply,
                     bs.part.as<double>(), (uint64_t)0, scal_out);
  GRE_HIP(hipGetLastError());
}

template <class M, typename TT>
void bpass_g(const BPlan& b, gre_pr* P, const M& m, double* scal_out, hipStream_t s, int sched) {
  switch (pow2ceil(P->k)) {
    case 1: bpass<1, M, TT>(b, P, m, scal_out, s, sched); break;
    case 2: bpass<2, M, TT>(b, P, m, scal_out, s, sched); break;
    case 4: bpass<4, M, TT>(b, P, m, scal_out, s, sched); break;
    case 8: bpass<8, M, TT>(b, P, m, scal_out, s, sched); break;
    case 16: bpass<16, M, TT>(b, P, m, scal_out, s, sched); break;
    case 32: bpass<32, M, TT>(b, P, m, scal_out, s, sched); break;
    default: bpass<64, M, TT>(b, P, m, scal_out, s, sched); break;
  }
}

// The batched schedule of direction d: the handle's fixed schedule, or (auto)
// the faster of the tile walk and the group-per-row kernel, timed on two
// passes after a warm-up one at first use.  pass(sched, it) runs one pass.
template <typename F>
int batch_sched(gre_pr* P, int d, hipStream_t s, F&& pass) {
  BatchState& bs = P->bstate;
  if (P->schedule == GRE_SCHED_ROWS) return GRE_SCHED_ROWS;
  if (P->schedule != GRE_SCHED_AUTO) return GRE_SCHED_TILE;
  if (bs.sched[d]) return bs.sched[d];
  hipEvent_t a, b;
  GRE_HIP(hipEventCreate(&a));
  GRE_HIP(hipEventCreate(&b));
  float best = 1e30f;
  int pick = GRE_SCHED_TILE;
  for (int sc : {GRE_SCHED_TILE, GRE_SCHED_ROWS}) {
    // the warm-up pass is timed too: a candidate whose warm-up alone takes
    // twice the best's two timed passes is dropped without timing them
    GRE_HIP(hipEventRecord(a, s));
    pass(sc, 0);
    GRE_HIP(hipEventRecord(b, s));
    GRE_HIP(hipEventSynchronize(b));
    float warm = 0.f;
    GRE_HIP(hipEventElapsedTime(&warm, a, b));
    if (warm > 2.f * best) continue;
    GRE_HIP(hipEventRecord(a, s));
    pass(sc, 1);
    pass(sc, 2);
    GRE_HIP(hipEventRecord(b, s));
    GRE_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    GRE_HIP(hipEventElapsedTime(&ms, a, b));
    if (ms < best) { best = ms; pick = sc; }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  bs.sched[d] = pick;
  return pick;
}

// Everything a captured batched run bakes into its launches.
std::vector<uint64_t> batch_key(const gre_pr* P, const BPlan& b, int sched, double damp, int iters,
                                std::initializer_list<const void*> ptrs) {
  uint64_t dbits = 0;
  std::memcpy(&dbits, &damp, sizeof(dbits));
  std::vector<uint64_t> k = {dbits, (uint64_t)iters, (uint64_t)P->dtype, (uint64_t)P->k, (uint64_t)sched,
                             b.ntiles, (uint64_t)b.nblk, (uint64_t)b.napply};
  for (const void* p : ptrs) k.push_back((uint64_t)(uintptr_t)p);
  const BatchState& bs = P->bstate;
  for (const void* p : {(const void*)b.off, (const void*)b.idx, (const void*)b.trb, (const void*)bs.carryP.p,
                        (const void*)bs.carryO.p, (const void*)bs.part.p, (const void*)bs.scal.p})
    k.push_back((uint64_t)(uintptr_t)p);
  return k;
}

// max over vectors of the L1 residual (scal layout: [a0_0, a1_0, a0_1, a1_1, ...])
double max_resid(const gre_pr* P, const double* dev, hipStream_t s) {
  std::vector<double> h(2 * P->k);
  GRE_HIP(hipMemcpyAsync(h.data(), dev, sizeof(double) * 2 * P->k, hipMemcpyDeviceToHost, s));
  GRE_HIP(hipStreamSynchronize(s));
  double m = 0.0;
  for (int j = 0; j < P->k; ++j) m = h[2 * j + 1] > m ? h[2 * j + 1] : m;
  return m;
}

template <typename TT>
double bforward(gre_pr* P, double damp, int iters, double tol, int* done) {
  hipStream_t s = stream();
  const BPlan b = make_plan(P, true);
  ensure_bscal(P, std::max(iters, 3));  // 3: the schedule timing passes
  BatchState& bs = P->bstate;
  const int k = P->k;
  double* sc = bs.scal.as<double>();
  const unsigned gb = gcap((uint64_t)P->g->nV * 64);
  double* tmp = static_cast<double*>(scratch(3, sizeof(double) * 2 * k * (size_t)gb));
  auto init = [&]() {
    P->xcur = 0;
    hipLaunchKernelGGL(k_binit<TT>, dim3(gb), dim3(kBlk), 0, s, P->g->nV, k, P->p.as<TT>(), P->p.as<TT>(),
                       P->g->outdeg.as<int32_t>(), P->x[0].as<TT>(), 0, tmp, (uint64_t)gb);
    hipLaunchKernelGGL(k_reduce_rows, dim3(2 * k), dim3(256), 0, s, tmp, (uint64_t)gb, tmp, (uint64_t)0, sc);
    GRE_HIP(hipGetLastError());
  };
  auto pass = [&](int sched, int it) {
    BFwd<TT> m;
    m.xin = P->x[P->xcur].as<TT>();
    m.xout = P->x[P->xcur ^ 1].as<TT>();
    m.w = P->wcsc.as<TT>();
    m.p = P->p.as<TT>();
    m.outdeg = P->g->outdeg.as<int32_t>();
    m.scal = sc + 2 * k * it;  // a0_j at index 2j
    m.damp = damp;
    m.k = k;
    m.resid = tol > 0.0 || it + 1 >= iters;
    bpass_g<BFwd<TT>, TT>(b, P, m, sc + 2 * k * (it + 1), s, sched);
    P->xcur ^= 1;
  };
  int sched = GRE_SCHED_TILE;
  if (P->schedule == GRE_SCHED_AUTO && !bs.sched[0]) init();  // the timing passes start from the initial state
  sched = batch_sched(P, 0, s, pass);
  if (tol <= 0.0 && iters > 0 && P->graphs && !prof_on()) {
    // fixed iteration count: one captured graph per direction, replayed while
    // nothing it depends on has changed
    const std::vector<uint64_t> key =
        batch_key(P, b, sched, damp, iters, {P->x[0].p, P->x[1].p, P->p.p, P->wcsc.p, tmp});
    if (!bs.exec[0] || key != bs.key[0]) {
      bs.drop(0);
      GRE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      init();
      for (int it = 0; it < iters; ++it) pass(sched, it);
      GRE_HIP(hipStreamEndCapture(s, &bs.graph[0]));
      GRE_HIP(hipGraphInstantiate(&bs.exec[0], bs.graph[0], nullptr, nullptr, 0));
      bs.key[0] = key;
      bs.post_cur[0] = P->xcur;
    }
    GRE_HIP(hipGraphLaunch(bs.exec[0], s));
    P->xcur = bs.post_cur[0];
    const double res = max_resid(P, sc + 2 * k * iters, s);
    if (done) *done = iters;
    return res;
  }
  init();
  int it;
  double res = 0.0;
  for (it = 0; it < iters; ++it) {
    pass(sched, it);
    if (tol > 0.0) {
      res = max_resid(P, sc + 2 * k * (it + 1), s);
      if (res < tol) { ++it; break; }
    }
  }
  if (tol <= 0.0 && iters > 0) res = max_resid(P, sc + 2 * k * iters, s);
  GRE_HIP(hipStreamSynchronize(s));
  if (done) *done = it;
  return res;
}

template <typename TT>
double badjoint(gre_pr* P, double damp, int iters, double tol, int* done) {
  hipStream_t s = stream();
  const BPlan b = make_plan(P, false);
  ensure_bscal(P, std::max(iters, 3));  // 3: the schedule timing passes
  BatchState& bs = P->bstate;
  const int k = P->k;
  double* sc = bs.scal.as<double>();
  const unsigned gb = gcap((uint64_t)P->g->nV * 64);
  double* tmp = static_cast<double*>(scratch(3, sizeof(double) * 2 * k * (size_t)gb));
  auto init = [&]() {
    P->lcur = 0;
    hipLaunchKernelGGL(k_binit<TT>, dim3(gb), dim3(kBlk), 0, s, P->g->nV, k, P->c.as<TT>(), P->p.as<TT>(),
                       P->g->outdeg.as<int32_t>(), P->lam[0].as<TT>(), 1, tmp, (uint64_t)gb);
    hipLaunchKernelGGL(k_reduce_rows, dim3(2 * k), dim3(256), 0, s, tmp, (uint64_t)gb, tmp, (uint64_t)0, sc);
    GRE_HIP(hipGetLastError());
  };
  auto pass = [&](int sched, int it) {
    BAdj<TT> m;
    m.lin = P->lam[P->lcur].as<TT>();
    m.lout = P->lam[P->lcur ^ 1].as<TT>();
    m.w = P->wcsr.as<TT>();
    m.c = P->c.as<TT>();
    m.p = P->p.as<TT>();
    m.outdeg = P->g->outdeg.as<int32_t>();
    m.wl = P->wl.as<TT>();
    m.scal = sc + 2 * k * it;
    m.damp = damp;
    m.k = k;
    m.resid = tol > 0.0 || it + 1 >= iters;
    bpass_g<BAdj<TT>, TT>(b, P, m, sc + 2 * k * (it + 1), s, sched);
    P->lcur ^= 1;
  };
  int sched = GRE_SCHED_TILE;
  if (P->schedule == GRE_SCHED_AUTO && !bs.sched[1]) init();  // the timing passes start from the initial state
  sched = batch_sched(P, 1, s, pass);
  if (tol <= 0.0 && iters > 0 && P->graphs && !prof_on()) {
    const std::vector<uint64_t> key = batch_key(
        P, b, sched, damp, iters, {P->lam[0].p, P->lam[1].p, P->c.p, P->p.p, P->wcsr.p, P->wl.p, tmp});
    if (!bs.exec[1] || key != bs.key[1]) {
      bs.drop(1);
      GRE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      init();
      for (int it = 0; it < iters; ++it) pass(sched, it);
      GRE_HIP(hipStreamEndCapture(s, &bs.graph[1]));
      GRE_HIP(hipGraphInstantiate(&bs.exec[1], bs.graph[1], nullptr, nullptr, 0));
      bs.key[1] = key;
      bs.post_cur[1] = P->lcur;
    }
    GRE_HIP(hipGraphLaunch(bs.exec[1], s));
    P->lcur = bs.post_cur[1];
    const double res = max_resid(P, sc + 2 * k * iters, s);
    if (done) *done = iters;
    return res;
  }
  init();
  int it;
  double res = 0.0;
  for (it = 0; it < iters; ++it) {
    pass(sched, it);
    if (tol > 0.0) {
      res = max_resid(P, sc + 2 * k * (it + 1), s);
      if (res < tol) { ++it; break; }
    }
  }
  if (tol <= 0.0 && iters > 0) res = max_resid(P, sc + 2 * k * iters, s);
  GRE_HIP(hipStreamSynchronize(s));
  if (done) *done = it;
  return res;
}

template <typename TT>
void bgrad(gre_pr* P, double damp, float* gW_h, float* graw_h, void* gp_h) {
  hipStream_t s = stream();
  gre_graph* g = P->g;
  const int k = P->k;
  const TT* x = P->x[P->xcur].as<TT>();
  const TT* lam = P->lam[P->lcur].as<TT>();
  const unsigned gb = gcap((uint64_t)g->nV * 64);
  double* tmp = static_cast<double*>(scratch(3, sizeof(double) * 2 * k * (size_t)gb));
  double* dang = static_cast<double*>(scratch(4, sizeof(double) * 2 * k));
  DevBuf tmpv;
  tmpv.alloc(tsz(P) * (size_t)g->nV * k);
  hipLaunchKernelGGL(k_binit<TT>, dim3(gb), dim3(kBlk), 0, s, g->nV, k, x, x, g->outdeg.as<int32_t>(), tmpv.as<TT>(), 0,
                     tmp, (uint64_t)gb);
  hipLaunchKernelGGL(k_reduce_rows, dim3(2 * k), dim3(256), 0, s, tmp, (uint64_t)gb, tmp, (uint64_t)0, dang);
  if (gp_h) {
    // dang_j sits at index 2j: pass a strided view via a compacting copy
    std::vector<double> h(2 * k), dj(k);
    GRE_HIP(hipMemcpyAsync(h.data(), dang, sizeof(double) * 2 * k, hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
    for (int j = 0; j < k; ++j) dj[j] = h[2 * j];
    GRE_HIP(hipMemcpyAsync(dang, dj.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_bgrad_p<TT>, dim3(gcap((uint64_t)g->nV * k)), dim3(kBlk), 0, s, g->nV, k, lam, dang, damp,
                       tmpv.as<TT>());
    GRE_HIP(hipMemcpyAsync(gp_h, tmpv.p, tsz(P) * (size_t)g->nV * k, hipMemcpyDeviceToHost, s));
  }
  if ((gW_h || graw_h) && g->nE > 0) {
    uint64_t nt = 0;
    const uint32_t* trb = graph_tiles(g, false, kTileG, &nt);
    DevBuf dgW, dgraw;
    dgW.alloc(4 * (size_t)g->nE);
    if (graw_h) dgraw.alloc(4 * (size_t)g->nE);
    const int G = pow2ceil(k);
    const unsigned nb = (unsigned)((nt + kBlk / G - 1) / (kBlk / G));
    {
      ProfScope ps("batch_edge_grad");
#define GRE_BGRAD(GG)                                                                                          \
  hipLaunchKernelGGL((k_bgrad<GG, TT>), dim3(nb), dim3(kBlk), 0, s, g->off.as<uint64_t>(), g->idx.as<uint32_t>(), \
                     trb, g->nE, g->nV, nt, k, P->rowsum.as<double>(), x, lam, P->wl.as<TT>(), damp,            \
                     dgW.as<float>(), graw_h ? dgraw.as<float>() : nullptr)
      switch (G) {
        case 1: GRE_BGRAD(1); break;
        case 2: GRE_BGRAD(2); break;
        case 4: GRE_BGRAD(4); break;
        case 8: GRE_BGRAD(8); break;
        case 16: GRE_BGRAD(16); break;
        case 32: GRE_BGRAD(32); break;
        default: GRE_BGRAD(64); break;
      }
#undef GRE_BGRAD
    }
    GRE_HIP(hipGetLastError());
    if (gW_h) GRE_HIP(hipMemcpyAsync(gW_h, dgW.p, 4 * g->nE, hipMemcpyDeviceToHost, s));
    if (graw_h) GRE_HIP(hipMemcpyAsync(graw_h, dgraw.p, 4 * g->nE, hipMemcpyDeviceToHost, s));
  }
  GRE_HIP(hipStreamSynchronize(s));
}

}  // namespace

double batched_forward(gre_pr* P, double damp, int iters, double tol, int* done) {
  return P->dtype == GRE_F64 ? bforward<double>(P, damp, iters, tol, done) : bforward<float>(P, damp, iters, tol, done);
}

double batched_adjoint(gre_pr* P, double damp, int iters, double tol, int* done) {
  return P->dtype == GRE_F64 ? badjoint<double>(P, damp, iters, tol, done) : badjoint<float>(P, damp, iters, tol, done);
}

void batched_targets(gre_pr* P, const uint32_t* targets_dev, int64_t nt) {
  hipStream_t s = stream();
  if (P->dtype == GRE_F64)
    hipLaunchKernelGGL(k_btargets<double>, dim3(gcap((uint64_t)nt * P->k)), dim3(kBlk), 0, s, targets_dev, nt, P->k,
                       P->c.as<double>());
  else
    hipLaunchKernelGGL(k_btargets<float>, dim3(gcap((uint64_t)nt * P->k)), dim3(kBlk), 0, s, targets_dev, nt, P->k,
                       P->c.as<float>());
  GRE_HIP(hipGetLastError());
}

void batched_gradients(gre_pr* P, double damp, float* gW, float* graw, void* gp) {
  if (P->dtype == GRE_F64) bgrad<double>(P, damp, gW, graw, gp);
  else bgrad<float>(P, damp, gW, graw, gp);
}

}  // namespace gre
