// PageRank / personalised PageRank forward, adjoint and edge gradients.
//
// Forward (pull over CSC):  x' = d (W^T x + (g.x) p) + (1-d) p
//   unweighted: W_uv = 1/outdeg(u), p uniform -> the classic power iteration
//   (1/N start, uniform teleport, dangling mass spread uniformly);
// Adjoint (pull over CSR):  lambda' = c + d (W lambda + g (p.lambda))
// Gradients:  dL/dW_uv = d x_u lambda_v,
//             dL/dw_uv = d x_u (lambda_v - (W lambda)_u) / rowsum_u,
//             dL/dp_v  = ((1-d) + d (g.x)) lambda_v.
// The FP64 reductions (dangling mass g.x, p.lambda, L1 residual) are fixed
// order block partials, so every run gives the same bits.
#include "pr.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>

// Device memory of the allocation-free forward (gre_pr_run_noalloc): the
// init pass's dangling-mass partials (two doubles per block, at most 65,536
// blocks), then each iteration's reduction pair (dangling mass, L1 change) for
// at most kNoAllocIters iterations: 1,114,160 bytes per GPU, held until the
// process ends.  Host code never passes this name to a kernel: it is a
// host-side stand-in, and noalloc_scratch() looks up each device's own address.
namespace gre {
namespace {
constexpr int kNoAllocIters = 4096;
constexpr size_t kNoAllocTmp = 2 * 65536;
constexpr size_t kNoAllocDoubles = kNoAllocTmp + 2 * (size_t)(kNoAllocIters + 1);
}  // namespace
}  // namespace gre
static __device__ __align__(256) double scratch[gre::kNoAllocDoubles];

namespace gre {
namespace {

constexpr unsigned kB = 256;

unsigned gcap(uint64_t n) {
  uint64_t g = (n + kB - 1) / kB;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

template <typename TT>
struct ModeFwd {
  using T = TT;
  const TT* gin;          // unweighted: contrib ; weighted: x
  TT* contrib_out;        // unweighted only
  const TT* xin;          // previous ranks (residual)
  TT* xout;               // new ranks
  const TT* w;            // normalised weights in CSC order (weighted), compute type
  const TT* p;            // personalisation or nullptr (uniform)
  const int32_t* outdeg;
  const uint32_t* cid;    // compact source ids (contrib indexed by cid[v]) or nullptr
  const double* scal_in;  // [0] = dangling mass of xin
  double damp, invn;
  bool resid = true;      // accumulate the L1 change (fixed-count runs need it on the last iteration only)
  bool craw = false;      // contrib_out takes the rank itself (weighted), not rank / out-degree
  // what row r's update reads besides its sum, loaded ahead by staged pulls
  // (R: with the old rank, for the L1 change)
  struct Pre {
    TT pv, old;
    int32_t d;
    uint32_t c;
  };
  template <bool R>
  __device__ Pre pre_r(uint32_t r) const {
    Pre q;
    q.pv = p ? p[r] : (TT)invn;
    q.old = R ? xin[r] : (TT)0;
    q.d = outdeg[r];
    q.c = (contrib_out && cid) ? cid[r] : 0u;
    return q;
  }
  template <bool R>
  __device__ void finalize_r(uint32_t r, TT s, const Pre& q, double& a0, double& a1) const {
    const double dang = scal_in[0];
    const TT y = (TT)damp * (s + (TT)dang * q.pv) + (TT)(1.0 - damp) * q.pv;
    xout[r] = y;
    const int32_t d = q.d;
    if (contrib_out) {
      if (cid) { if (d > 0) contrib_out[q.c] = craw ? y : y / (TT)d; }
      else contrib_out[r] = craw ? y : (d > 0 ? y / (TT)d : (TT)0);
    }
    if (d == 0) a0 += (double)y;
    if (R) a1 += fabs((double)y - (double)q.old);
  }
};

// The weighting and the L1 change are compile-time properties of the launched
// mode: a run-time test inside gather() would put each weighted gather in its
// own branch, where the multiply waits for the loads before the next gather is
// issued, and a conditional old-rank load slows the row pass the same way.
template <typename TT, bool W, bool R>
struct ModeFwdG : ModeFwd<TT> {
  using Pre = typename ModeFwd<TT>::Pre;
  __device__ TT gather(uint64_t e, uint32_t u) const {
    if (W) return (TT)this->w[e] * this->gin[u];
    return this->gin[u];
  }
  __device__ Pre pre(uint32_t r) const { return this->template pre_r<R>(r); }
  __device__ void finalize_pre(uint32_t r, TT s, const Pre& q, double& a0, double& a1) const {
    this->template finalize_r<R>(r, s, q, a0, a1);
  }
  __device__ void finalize(uint32_t r, TT s, double& a0, double& a1) const { finalize_pre(r, s, pre(r), a0, a1); }
};
template <typename TT>
void fwd_pull(const PullPlan& p, const ModeFwd<TT>& m, double* scal_out, hipStream_t s) {
  if (m.w) {
    if (m.resid) pull_pass(p, ModeFwdG<TT, true, true>{m}, scal_out, s);
    else pull_pass(p, ModeFwdG<TT, true, false>{m}, scal_out, s);
  } else {
    if (m.resid) pull_pass(p, ModeFwdG<TT, false, true>{m}, scal_out, s);
    else pull_pass(p, ModeFwdG<TT, false, false>{m}, scal_out, s);
  }
}

template <typename TT>
struct ModeAdj {
  using T = TT;
  const TT* lin;          // previous lambda (residual)
  TT* lout;
  const TT* gin;          // gathered lambda: lin, or its compact-target-order copy
  TT* lc_out;             // compact schedules: lout in compact target order, else nullptr
  const uint32_t* tid;    // compact target ids (lc_out indexed by tid[u])
  const TT* w;            // normalised weights, CSR order (weighted and gathered here), compute type
  bool unw;               // unweighted: the sum is divided by the out-degree
  const TT* c;
  const TT* p;
  const int32_t* outdeg;
  TT* wl;                 // (W lambda)_u
  const double* scal_in;  // [0] = p.lambda of lin
  double damp, invn;
  // what row u's update reads besides its sum, loaded ahead by the tile pull
  struct Pre {
    TT cu, old, pv;
    int32_t d;
    uint32_t t;
  };
  __device__ Pre pre(uint32_t u) const {
    Pre q;
    q.t = lc_out ? tid[u] : 0xffffffffu;
    q.d = outdeg[u];
    q.cu = c[u];
    q.old = lin[u];
    q.pv = p ? p[u] : (TT)invn;
    return q;
  }
  __device__ void finalize_pre(uint32_t u, TT s, const Pre& q, double& a0, double& a1) const {
    const int32_t d = q.d;
    if (unw && d > 0) s = s / (TT)d;
    wl[u] = s;
    const TT y = q.cu + (TT)damp * (s + (d == 0 ? (TT)scal_in[0] : (TT)0));
    lout[u] = y;
    if (q.t != 0xffffffffu) lc_out[q.t] = y;
    a0 += (double)q.pv * (double)y;
    a1 += fabs((double)y - (double)q.old);
  }
  __device__ void finalize(uint32_t u, TT s, double& a0, double& a1) const { finalize_pre(u, s, pre(u), a0, a1); }
};

template <typename TT, bool W>
struct ModeAdjG : ModeAdj<TT> {
  __device__ TT gather(uint64_t e, uint32_t v) const {
    if (W) return (TT)this->w[e] * this->gin[v];
    return this->gin[v];
  }
};
template <typename TT>
void adj_pull(const PullPlan& p, const ModeAdj<TT>& m, double* scal_out, hipStream_t s) {
  if (m.w) pull_pass(p, ModeAdjG<TT, true>{m}, scal_out, s);
  else pull_pass(p, ModeAdjG<TT, false>{m}, scal_out, s);
}

// partitioned pull, first pass: a segment's sum of contrib over its sources
template <typename TT, bool W>
struct ModeSeg {
  using T = TT;
  const TT* gin;  // contrib, compact source ids
  TT* sum;        // per segment
  const TT* w;    // W: normalised weights in the partition's entry order
  __device__ TT gather(uint64_t e, uint32_t u) const {
    if (W) return w[e] * gin[u];
    return gin[u];
  }
  __device__ void finalize(uint32_t seg, TT s, double&, double&) const { sum[seg] = s; }
};

// row sums of the raw weights (pull over CSR, gather = weight)
struct ModeRowSum {
  using T = double;
  const float* w;
  double* rowsum;
  __device__ double gather(uint64_t e, uint32_t) const { return (double)w[e]; }
  __device__ void finalize(uint32_t r, double s, double&, double&) const { rowsum[r] = s; }
};

template <typename TT>
__global__ void k_fwd_init(uint32_t nV, const TT* __restrict__ p, const int32_t* __restrict__ outdeg,
                           TT* __restrict__ x, TT* __restrict__ contrib, double invn, double* __restrict__ part,
                           const uint32_t* __restrict__ cid, bool craw) {
  __shared__ double red[8];
  double a0 = 0.0, a1 = 0.0;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const TT xv = p ? p[v] : (TT)invn;
    x[v] = xv;
    const int32_t d = outdeg[v];
    if (contrib) {
      if (cid) { if (d > 0) contrib[cid[v]] = craw ? xv : xv / (TT)d; }
      else contrib[v] = craw ? xv : (d > 0 ? xv / (TT)d : (TT)0);
    }
    if (d == 0) a0 += (double)xv;
  }
  block_sum2(a0, a1, red);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = a0; part[2 * blockIdx.x + 1] = a1; }
}

// The same over relabelled partition rows: everything in row order (p_r,
// outdeg_r, cid_r; x written at the row), contrib by compact id as before.
template <typename TT>
__global__ void k_fwd_init_rows(uint32_t nV, const TT* __restrict__ p_r, const int32_t* __restrict__ outdeg_r,
                                TT* __restrict__ x, TT* __restrict__ contrib, double invn,
                                double* __restrict__ part, const uint32_t* __restrict__ cid_r, bool craw) {
  __shared__ double red[8];
  double a0 = 0.0, a1 = 0.0;
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < nV;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const TT xv = p_r ? p_r[r] : (TT)invn;
    x[r] = xv;
    const int32_t d = outdeg_r[r];
    if (d > 0) contrib[cid_r[r]] = craw ? xv : xv / (TT)d;
    if (d == 0) a0 += (double)xv;
  }
  block_sum2(a0, a1, red);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = a0; part[2 * blockIdx.x + 1] = a1; }
}

// a vertex-order array into the partition's row order: out[rowperm[v]] = in[v]
template <typename TT>
__global__ void k_vertices_to_rows(uint32_t nV, const uint32_t* __restrict__ rowperm, const TT* __restrict__ in,
                                   TT* __restrict__ out) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV; v += (uint64_t)gridDim.x * blockDim.x)
    out[rowperm[v]] = in[v];
}

// the ranks back from the partition's row order: out[v] = in[rowperm[v]]
template <typename TT>
__global__ void k_rows_to_vertices(uint32_t nV, const uint32_t* __restrict__ rowperm, const TT* __restrict__ in,
                                   TT* __restrict__ out) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV; v += (uint64_t)gridDim.x * blockDim.x)
    out[v] = in[rowperm[v]];
}

template <typename TT>
__global__ void k_adj_init(uint32_t nV, const TT* __restrict__ c, const TT* __restrict__ p, TT* __restrict__ lam,
                           double invn, double* __restrict__ part, const uint32_t* __restrict__ tid,
                           TT* __restrict__ lc) {
  __shared__ double red[8];
  double a0 = 0.0, a1 = 0.0;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV;
       v += (uint64_t)gridDim.x * blockDim.x) {
    lam[v] = c[v];
    if (lc) {
      const uint32_t t = tid[v];
      if (t != 0xffffffffu) lc[t] = c[v];
    }
    a0 += (double)(p ? p[v] : (TT)invn) * (double)c[v];
  }
  block_sum2(a0, a1, red);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = a0; part[2 * blockIdx.x + 1] = a1; }
}

__global__ void k_lognormal(const uint64_t* __restrict__ off, const uint32_t* __restrict__ idx, uint32_t nV,
                            uint64_t nE, uint64_t seed, float* __restrict__ w) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nE;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = row_of(off, nV, e), v = idx[e];
    const uint64_t h1 = mix64(seed * GRE_K0 + (uint64_t)u * GRE_K1 + (uint64_t)v * GRE_K2 + GRE_K3);
    const uint64_t h2 = mix64(h1 ^ GRE_K4);
    const double u1 = ((double)(h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
    const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    w[e] = (float)exp(z);
  }
}

template <typename TT>
__global__ void k_normalise(const uint64_t* __restrict__ off, uint32_t nV, uint64_t nE, const float* __restrict__ w,
                            const double* __restrict__ rowsum, TT* __restrict__ wn) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nE;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = row_of(off, nV, e);
    wn[e] = (TT)((double)w[e] / rowsum[u]);
  }
}

template <typename TT>
__global__ void k_permute(const TT* __restrict__ in, const uint64_t* __restrict__ perm, uint64_t n,
                          TT* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = in[perm[i]];
}

template <typename TT>
__global__ void k_permute32(const TT* __restrict__ in, const uint32_t* __restrict__ perm, uint64_t n,
                            TT* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = in[perm[i]];
}

template <typename TT>
__global__ void k_targets(uint32_t nV, const uint32_t* __restrict__ tg, int64_t nt, TT* __restrict__ c) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nt; i += (int64_t)gridDim.x * blockDim.x)
    c[tg[i]] = (TT)1;
}

// gW/graw over CSR; one block per CSR tile, row found inside the tile's range
template <typename TT>
__global__ __launch_bounds__(256) void k_edge_grad(const uint64_t* __restrict__ off, const uint32_t* __restrict__ idx,
                                                   const uint32_t* __restrict__ tile_rb, uint64_t E, uint32_t V,
                                                   uint64_t ntiles, uint32_t te, const void* __restrict__ w,
                                                   const double* __restrict__ rowsum, const TT* __restrict__ x,
                                                   const TT* __restrict__ lam, const TT* __restrict__ wl,
                                                   const int32_t* __restrict__ outdeg, double damp,
                                                   float* __restrict__ gW, float* __restrict__ graw) {
  const uint64_t t = blockIdx.x;
  const uint64_t e0 = t * (uint64_t)te;
  const uint64_t e1 = e0 + te < E ? e0 + te : E;
  uint32_t lo0 = tile_rb[t];
  lo0 = lo0 > 0 ? lo0 - 1 : 0;
  const uint32_t hi0 = (t + 1 == ntiles) ? V : tile_rb[t + 1];
  for (uint64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    uint32_t lo = lo0, hi = hi0 < V ? hi0 + 1 : V;  // answer in [lo, hi)
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (off[mid] <= e) lo = mid; else hi = mid;
    }
    const uint32_t u = lo, v = idx[e];
    const double xu = (double)x[u];
    const double lv = (double)lam[v];
    gW[e] = (float)(damp * xu * lv);
    if (graw) {
      const double rs = w ? rowsum[u] : (double)outdeg[u];
      graw[e] = rs > 0.0 ? (float)(damp * xu * (lv - (double)wl[u]) / rs) : 0.f;
    }
  }
}

template <typename TT>
__global__ void k_grad_p(uint32_t nV, const TT* __restrict__ lam, const double* __restrict__ dang, double damp,
                         TT* __restrict__ gp) {
  const double f = (1.0 - damp) + damp * dang[0];
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV;
       v += (uint64_t)gridDim.x * blockDim.x)
    gp[v] = (TT)(f * (double)lam[v]);
}

size_t tsize(int dtype) { return dtype == GRE_F64 ? 8 : 4; }

void ensure_scal(gre_pr* P, int iters) {
  if (P->scal_cap >= iters + 2) return;
  P->scal_cap = iters + 2;
  P->scal.alloc(sizeof(double) * 2 * (size_t)P->scal_cap);
}

// The compact schedules apply to the single-vector forward: its gathers read
// contrib[], which then holds one entry per vertex with out-edges, hubs first
// (graph_compact_sources); weighted runs multiply by the edge's weight, in CSC
// order (compact) or in the partition's entry order (part).
bool use_compact(const gre_pr* P) { return P->compact && P->k == 1; }
bool use_part(const gre_pr* P) { return P->part && use_compact(P); }
// the binned pull: single vector, FP32
bool use_bin(const gre_pr* P) { return P->bin && P->k == 1; }
// The single-vector adjoint's schedule family: the compact and partitioned
// forwards' own; the binned forward's adjoint runs its layout over the
// reversed graph (adj_bin), else the partitioned adjoint over the CSR with the
// handle's (ranges, shared).
bool adj_part(const gre_pr* P) {
  return use_part(P) || (use_bin(P) && P->g->nE > 0 && P->g->nE < 0xffffffffULL);
}
bool adj_compact(const gre_pr* P) { return use_compact(P) || use_bin(P); }
// The binned forward's adjoint, FP32: the binned layout built over the reversed
// graph, when its rows (the vertices with out-edges) fit the bins
bool adj_bin(gre_pr* P) {
  if (!use_bin(P) || P->dtype != GRE_F32 || P->g->nE == 0 || P->g->nE >= 0xffffffffULL) return false;
  gre_graph* r = graph_reversed(P->g);
  graph_compact_targets(r);
  return (uint64_t)r->ntgt <= (uint64_t)kBinMaxBins * kBinRows;
}
// the partitioned and binned forwards walk relabelled rows: a run keeps its
// ranks in the layout's row order until ranks_to_vertices
bool rows_relabelled(const gre_pr* P) { return (use_part(P) || use_bin(P)) && P->xp && P->xp->permuted; }

// Why the binned pull cannot run this handle, or nullptr.
const char* bin_refusal(const gre_pr* P) {
  if (P->k != 1) return "the binned schedule runs single-vector handles only";
  if (P->dtype != GRE_F32) return "the binned schedule runs FP32 handles only";
  if (P->g->nE == 0 || P->g->nE >= 0xffffffffULL) return "the binned schedule needs 0 < edges < 2^32";
  if (P->has_p && P->p_l1 > kBinMaxPL1)
    return "the binned schedule needs a personalisation of L1 norm at most 1 (its fixed-point sums are bounded by it)";
  return nullptr;
}

// the binned layout and the handle's buffers for it
void ensure_bin_plan(gre_pr* P) {
  if (const char* why = bin_refusal(P)) throw Error(GRE_DEV_NULL, why);
  if (!P->xp || !P->xp->binned) {
    P->xp.reset();
    P->wpart_xp.reset();
    P->wpart_ver = ~0ULL;
    P->plan_ipt_x = 0;
    P->xp = graph_binpart(P->g);
    if (!P->tuning) prune_xparts(P->g);
  }
  const size_t xb = sizeof(float) * (size_t)P->g->nV;
  if (P->x[1].bytes < xb) P->x[1].alloc(xb);
  if (P->p_r.bytes < xb) P->p_r.alloc(xb);
  if (P->bbuf.for_xp != P->xp.get()) P->bbuf.make(*P->xp, P->g->nV, stream());
  // weighted: the normalised weights in the layout's word order
  if (P->weighted && (P->wpart_xp != P->xp || P->wpart_ver != P->wver)) {
    const uint64_t Ew = P->xp->bin.Ew;
    hipStream_t s = stream();
    P->wpart.alloc(sizeof(float) * Ew);
    hipLaunchKernelGGL(k_permute32<float>, dim3(gcap(Ew)), dim3(kB), 0, s, P->wcsr.as<float>(),
                       P->xp->bin.wpos.as<uint32_t>(), Ew, P->wpart.as<float>());
    GRE_HIP(hipGetLastError());
    P->wpart_xp = P->xp;
    P->wpart_ver = P->wver;
  }
}

// the source-range partition, its segment plan and the segment sums
void ensure_part_plan(gre_pr* P, int ipt) {
  if (!P->xp || P->xp->R != P->ranges || P->xp->H != std::min(P->shared, P->g->nsrc)) {
    // let go of the previous partition (and the weight copy that follows it),
    // take the new one, then drop the old one unless someone else holds it
    P->xp.reset();
    P->wpart_xp.reset();
    P->wpart_ver = ~0ULL;
    P->xp = graph_xpart(P->g, P->ranges, P->shared, true, true);
    if (!P->tuning) prune_xparts(P->g);
    P->plan_ipt_x = 0;
  }
  // a run keeps its ranks (and the personalisation) in the partition's row
  // order and turns them back into the other rank buffer at its end
  const size_t xb = tsize(P->dtype) * (size_t)P->g->nV;
  if (P->x[1].bytes < xb) P->x[1].alloc(xb);
  if (P->p_r.bytes < xb) P->p_r.alloc(xb);
  if (P->plan_ipt_x != ipt) {
    P->buf_x.make(P->plan_x, *P->xp, ipt, tsize(P->dtype));
    // the per-row pass always uses 8 entries per thread
    P->buf_px.make_pairs(P->plan_px, *P->xp, P->g->nV, 8, tsize(P->dtype));
    P->plan_px.tag = "part_combine";
    P->plan_px.fix_tag = "part_combine_fixup";
    P->plan_ipt_x = ipt;
  }
  P->plan_x.nt = P->nt;
  P->plan_px.nt = false;
  const size_t need = tsize(P->dtype) * (size_t)P->xp->nseg;
  if (P->xsum.bytes < need) P->xsum.alloc(need);
  // weighted: the normalised weights follow the partition's entry order
  if (P->weighted && (P->wpart_xp != P->xp || P->wpart_ver != P->wver)) {
    const uint64_t nE = P->g->nE;
    hipStream_t s = stream();
    P->wpart.alloc(tsize(P->dtype) * nE);
    if (nE > 0) {
      if (P->dtype == GRE_F64)
        hipLaunchKernelGGL(k_permute32<double>, dim3(gcap(nE)), dim3(kB), 0, s, P->wcsc.as<double>(),
                           P->xp->perm.as<uint32_t>(), nE, P->wpart.as<double>());
      else
        hipLaunchKernelGGL(k_permute32<float>, dim3(gcap(nE)), dim3(kB), 0, s, P->wcsc.as<float>(),
                           P->xp->perm.as<uint32_t>(), nE, P->wpart.as<float>());
      GRE_HIP(hipGetLastError());
    }
    P->wpart_xp = P->xp;
    P->wpart_ver = P->wver;
  }
}

// the adjoint's partition: the same (R, H) over the CSR in compact target ids
void ensure_part_plan_adj(gre_pr* P, int ipt) {
  if (!P->xpa || P->xpa->R != P->ranges || P->xpa->H != std::min(P->shared, P->g->ntgt)) {
    P->xpa.reset();
    P->wparta_xp.reset();
    P->wparta_ver = ~0ULL;
    P->xpa = graph_xpart(P->g, P->ranges, P->shared, false);
    prune_xparts(P->g);
    P->plan_ipt_xa = 0;
  }
  if (P->plan_ipt_xa != ipt) {
    P->buf_xa.make(P->plan_xa, *P->xpa, ipt, tsize(P->dtype));
    P->buf_pxa.make_pairs(P->plan_pxa, *P->xpa, P->g->nV, 8, tsize(P->dtype));
    P->plan_xa.tag = "adj_part_segments";
    P->plan_pxa.tag = "adj_part_combine";
    P->plan_pxa.fix_tag = "adj_part_combine_fixup";
    P->plan_ipt_xa = ipt;
  }
  P->plan_xa.nt = P->nt;
  P->plan_pxa.nt = false;
  const size_t need = tsize(P->dtype) * (size_t)P->xpa->nseg;
  if (P->asum.bytes < need) P->asum.alloc(need);
  if (P->weighted && (P->wparta_xp != P->xpa || P->wparta_ver != P->wver)) {
    const uint64_t nE = P->g->nE;
    hipStream_t s = stream();
    P->wparta.alloc(tsize(P->dtype) * nE);
    if (nE > 0) {
      if (P->dtype == GRE_F64)
        hipLaunchKernelGGL(k_permute32<double>, dim3(gcap(nE)), dim3(kB), 0, s, P->wcsr.as<double>(),
                           P->xpa->perm.as<uint32_t>(), nE, P->wparta.as<double>());
      else
        hipLaunchKernelGGL(k_permute32<float>, dim3(gcap(nE)), dim3(kB), 0, s, P->wcsr.as<float>(),
                           P->xpa->perm.as<uint32_t>(), nE, P->wparta.as<float>());
      GRE_HIP(hipGetLastError());
    }
    P->wparta_xp = P->xpa;
    P->wparta_ver = P->wver;
  }
}

void ensure_plan(gre_pr* P, bool csc, int ipt) {
  if (csc && use_bin(P)) {
    ensure_bin_plan(P);
    return;
  }
  if (csc && use_part(P)) {
    graph_compact_sources(P->g);
    ensure_part_plan(P, ipt);
    return;
  }
  if (!csc && adj_part(P)) {
    graph_compact_targets(P->g);
    ensure_part_plan_adj(P, ipt);
    return;
  }
  int& have = csc ? P->plan_ipt_csc : P->plan_ipt_csr;
  PullPlan& plan = csc ? P->plan_csc : P->plan_csr;
  if (have != ipt) {
    (csc ? P->buf_csc : P->buf_csr).make(plan, P->g, csc, ipt, tsize(P->dtype));
    have = ipt;
  }
  plan.nt = P->nt;
  plan.idx = (csc ? P->g->cidx : P->g->idx).as<uint32_t>();
  if (csc && use_compact(P)) {
    graph_compact_sources(P->g);
    plan.idx = P->g->cidx_c.as<uint32_t>();
  }
  // the adjoint of the compact schedules gathers lambda by compact target id
  if (!csc && adj_compact(P)) {
    graph_compact_targets(P->g);
    plan.idx = P->g->idx_c.as<uint32_t>();
  }
}

template <typename TT>
ModeFwd<TT> fwd_mode(gre_pr* P, double damp) {
  // weighted: ranks ping-pong between x[0] and x[1] because they are gathered;
  // unweighted: ranks stay in x[0] (only their owner row touches them) and the
  // gathered x/outdeg ping-pongs between contrib[0] and contrib[1]
  const int xc = P->xcur, cc = P->ccur;
  ModeFwd<TT> m;
  // weighted and not compact: the ranks themselves are gathered
  const bool direct = P->weighted && !use_compact(P);
  m.w = P->weighted ? P->wcsc.as<TT>() : nullptr;
  m.gin = direct ? P->x[xc].as<TT>() : P->contrib[cc].as<TT>();
  m.contrib_out = direct ? nullptr : P->contrib[cc ^ 1].as<TT>();
  m.craw = P->weighted;
  m.xin = P->x[xc].as<TT>();
  m.xout = P->weighted ? P->x[xc ^ 1].as<TT>() : P->x[xc].as<TT>();
  m.p = P->has_p ? P->p.as<TT>() : nullptr;
  m.outdeg = P->g->outdeg.as<int32_t>();
  m.cid = use_compact(P) ? P->g->cid.as<uint32_t>() : nullptr;
  if (rows_relabelled(P)) {
    // the row pass walks the partition's rows: their data in row order
    m.p = P->has_p ? P->p_r.as<TT>() : nullptr;
    m.outdeg = P->xp->outdeg_r.as<int32_t>();
    m.cid = P->xp->cid_r.as<uint32_t>();
  }
  m.damp = damp;
  m.invn = 1.0 / (double)P->g->nV;
  return m;
}

// one forward iteration; scal_in/scal_out are device pointers to (acc0, acc1)
template <typename TT>
void fwd_iter(gre_pr* P, double damp, const double* scal_in, double* scal_out, hipStream_t s, bool resid = true) {
  ModeFwd<TT> m = fwd_mode<TT>(P, damp);
  m.scal_in = scal_in;
  m.resid = resid;
  if constexpr (std::is_same<TT, float>::value) {
    if (use_bin(P)) {
      BinIter a;
      a.wbin = P->weighted ? P->wpart.as<float>() : nullptr;
      a.contrib_in = P->contrib[P->ccur].as<float>();
      a.contrib_out = P->contrib[P->ccur ^ 1].as<float>();
      a.x = P->x[P->xcur].as<float>();
      a.p_r = P->has_p ? P->p_r.as<float>() : nullptr;
      a.outdeg_r = P->xp->outdeg_r.as<int32_t>();
      a.scal_in = scal_in;
      a.scal_out = scal_out;
      a.damp = damp;
      a.invn = 1.0 / (double)P->g->nV;
      a.resid = resid;
      bin_forward_iter(*P->xp, P->bbuf, a, P->g->nV, s);
      P->ccur ^= 1;
      return;
    }
  }
  if (use_part(P)) {
    // segment sums over XCD-local source ranges, then each row's segments
    if (P->weighted) pull_pass(P->plan_x, ModeSeg<TT, true>{m.gin, P->xsum.as<TT>(), P->wpart.as<TT>()}, nullptr, s);
    else pull_pass(P->plan_x, ModeSeg<TT, false>{m.gin, P->xsum.as<TT>(), nullptr}, nullptr, s);
    m.gin = P->xsum.as<TT>();
    m.w = nullptr;  // the row pass adds segment sums, weights already applied
    fwd_pull(P->plan_px, m, scal_out, s);
    if (P->weighted) P->xcur ^= 1;
    P->ccur ^= 1;
    return;
  }
  // in-place x for the unweighted form: residual reads xin before xout writes
  fwd_pull(P->plan_csc, m, scal_out, s);
  if (P->weighted) P->xcur ^= 1;
  if (m.contrib_out) P->ccur ^= 1;
}

// Dangling-mass partials of the init pass: two doubles per block.
size_t fwd_init_tmp_bytes(uint32_t nV) { return sizeof(double) * 2 * gcap(nV); }

// The forward's initial state; its dangling mass goes to scal[0..1] through
// the block partials in tmp (fwd_init_tmp_bytes).
template <typename TT>
void fwd_init(gre_pr* P, hipStream_t s, double* tmp, double* scal) {
  const uint32_t nV = P->g->nV;
  const unsigned gb = gcap(nV);
  P->xcur = 0;
  P->ccur = 0;
  if (rows_relabelled(P))
    hipLaunchKernelGGL(k_fwd_init_rows<TT>, dim3(gb), dim3(kB), 0, s, nV, P->has_p ? P->p_r.as<TT>() : nullptr,
                       P->xp->outdeg_r.as<int32_t>(), P->x[0].as<TT>(), P->contrib[0].as<TT>(), 1.0 / (double)nV,
                       tmp, P->xp->cid_r.as<uint32_t>(), P->weighted);
  else
    hipLaunchKernelGGL(k_fwd_init<TT>, dim3(gb), dim3(kB), 0, s, nV, P->has_p ? P->p.as<TT>() : nullptr,
                       P->g->outdeg.as<int32_t>(), P->x[0].as<TT>(),
                       (P->weighted && !use_compact(P)) ? nullptr : P->contrib[0].as<TT>(), 1.0 / (double)nV, tmp,
                       use_compact(P) ? P->g->cid.as<uint32_t>() : nullptr, P->weighted);
  hipLaunchKernelGGL(k_reduce_pairs, dim3(1), dim3(1024), 0, s, tmp, (uint64_t)gb, tmp, (uint64_t)0, scal);
  GRE_HIP(hipGetLastError());
}

// Before a run over relabelled partition rows: the personalisation in row
// order, if it changed since (outside any captured graph).
template <typename TT>
void fwd_prepare(gre_pr* P, hipStream_t s) {
  if (!rows_relabelled(P) || !P->has_p || (P->p_r_ver == P->pver && P->p_r_xp == P->xp)) return;
  const uint32_t nV = P->g->nV;
  hipLaunchKernelGGL(k_vertices_to_rows<TT>, dim3(gcap(nV)), dim3(kB), 0, s, nV, P->xp->rowperm.as<uint32_t>(),
                     P->p.as<TT>(), P->p_r.as<TT>());
  GRE_HIP(hipGetLastError());
  P->p_r_ver = P->pver;
  P->p_r_xp = P->xp;
}

// After a run: over relabelled partition rows the ranks are in row order
// until something reads them (ranks_to_vertices).
void fwd_done(gre_pr* P) {
  P->x_rows = rows_relabelled(P);
  P->rows_xp = P->x_rows ? P->xp : nullptr;
}

// The ranks back in vertex order, in the other rank buffer.
template <typename TT>
void ranks_to_vertices_t(gre_pr* P, hipStream_t s) {
  const uint32_t nV = P->g->nV;
  hipLaunchKernelGGL(k_rows_to_vertices<TT>, dim3(gcap(nV)), dim3(kB), 0, s, nV, P->rows_xp->rowperm.as<uint32_t>(),
                     P->x[P->xcur].as<TT>(), P->x[P->xcur ^ 1].as<TT>());
  GRE_HIP(hipGetLastError());
}
void ranks_to_vertices(gre_pr* P) {
  if (!P->x_rows) return;
  if (P->dtype == GRE_F64) ranks_to_vertices_t<double>(P, stream());
  else ranks_to_vertices_t<float>(P, stream());
  P->xcur ^= 1;
  P->x_rows = false;
  P->rows_xp.reset();
}

// Everything a captured forward run bakes into its kernels' arguments and
// grids: a replay is valid only while all of it is unchanged.  `tmp` is the
// shared reduction scratch the init pass writes: another handle may grow (and
// so move) it between replays.
std::vector<uint64_t> fwd_graph_key(const gre_pr* P, double damp, int iters, const void* tmp) {
  uint64_t dbits = 0;
  std::memcpy(&dbits, &damp, sizeof(dbits));
  auto ptr = [](const void* p) { return (uint64_t)(uintptr_t)p; };
  std::vector<uint64_t> k = {dbits, (uint64_t)iters, (uint64_t)P->dtype, (uint64_t)P->schedule, (uint64_t)P->ipt,
                             (uint64_t)P->nt, (uint64_t)P->compact, (uint64_t)P->part, (uint64_t)P->weighted,
                             (uint64_t)P->has_p, ptr(P->p.p), ptr(P->scal.p), ptr(P->x[0].p), ptr(P->x[1].p),
                             ptr(P->contrib[0].p), ptr(P->contrib[1].p), ptr(P->wcsc.p), ptr(P->xsum.p),
                             ptr(P->wpart.p), P->wver, ptr(tmp)};
  for (const PullPlan* pl : {&P->plan_csc, &P->plan_x, &P->plan_px}) {
    k.push_back(ptr(pl->off));
    k.push_back(ptr(pl->off32));
    k.push_back(ptr(pl->idx));
    k.push_back(ptr(pl->tile_rb));
    k.push_back(ptr(pl->carryP));
    k.push_back(ptr(pl->part));
    k.push_back(pl->ntiles);
    k.push_back((uint64_t)pl->ipt);
    k.push_back(ptr(pl->xmap));
    k.push_back(ptr(pl->slot));
    k.push_back(pl->grid);
  }
  k.push_back((uint64_t)P->bin);
  k.push_back(ptr(P->bbuf.binval.p));
  k.push_back(ptr(P->bbuf.gacc.p));
  k.push_back(ptr(P->bbuf.part.p));
  k.push_back(ptr(P->bbuf.part2.p));
  k.push_back(ptr(P->bbuf.part3.p));
  if (P->xp) {
    k.push_back((uint64_t)P->xp->binned);
    k.push_back(ptr(P->xp->bin.src16.p));
    k.push_back(ptr(P->xp->bin.dloc.p));
    k.push_back((uint64_t)P->xp->R);
    k.push_back(P->xp->H);
    k.push_back(P->xp->nseg);
    k.push_back(ptr(P->xp->rowperm.p));
    k.push_back(ptr(P->xp->outdeg_r.p));
    k.push_back(ptr(P->xp->cid_r.p));
  }
  k.push_back(ptr(P->p_r.p));
  return k;
}

template <typename TT>
double forward_eager(gre_pr* P, double damp, int iters, double tol, int* done, double* tmp, double* sc, bool sync);

template <typename TT>
double run_forward(gre_pr* P, double damp, int iters, double tol, int* done) {
  hipStream_t s = stream();
  if (tol <= 0.0 && iters > 0 && P->graphs && !prof_on()) {
    // fixed iteration count: the whole run is one captured graph, replayed
    // while nothing it depends on has changed
    // the scratch is sized (and so placed) before the key is taken
    double* tmp = static_cast<double*>(scratch(3, fwd_init_tmp_bytes(P->g->nV)));
    std::vector<uint64_t> key = fwd_graph_key(P, damp, iters, tmp);
    fwd_prepare<TT>(P, s);
    if (!P->fwd_exec || key != P->fwd_key) {
      if (P->fwd_exec) { (void)hipGraphExecDestroy(P->fwd_exec); P->fwd_exec = nullptr; }
      if (P->fwd_graph) { (void)hipGraphDestroy(P->fwd_graph); P->fwd_graph = nullptr; }
      GRE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      double* sc = P->scal.as<double>();
      fwd_init<TT>(P, s, tmp, sc);
      // only the last iteration's L1 change is returned
      for (int it = 0; it < iters; ++it) fwd_iter<TT>(P, damp, sc + 2 * it, sc + 2 * (it + 1), s, it + 1 == iters);
      GRE_HIP(hipStreamEndCapture(s, &P->fwd_graph));
      GRE_HIP(hipGraphInstantiate(&P->fwd_exec, P->fwd_graph, nullptr, nullptr, 0));
      P->fwd_key = key;
      P->fwd_post_xcur = P->xcur;
      P->fwd_post_ccur = P->ccur;
    }
    GRE_HIP(hipGraphLaunch(P->fwd_exec, s));
    P->xcur = P->fwd_post_xcur;
    P->ccur = P->fwd_post_ccur;
    fwd_done(P);
    double h[2];
    GRE_HIP(hipMemcpyAsync(h, P->scal.as<double>() + 2 * iters, sizeof(h), hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
    if (done) *done = iters;
    return h[1];
  }
  return forward_eager<TT>(P, damp, iters, tol, done, static_cast<double*>(scratch(3, fwd_init_tmp_bytes(P->g->nV))),
                           P->scal.as<double>(), true);
}

// The forward launched kernel by kernel, with its reductions in tmp (the init
// pass's partials) and sc (2 * (iters + 1) doubles: each iteration's dangling
// mass and L1 change).  sync = false with tol <= 0 only enqueues the work and
// returns 0 (no host wait, no read-back).
template <typename TT>
double forward_eager(gre_pr* P, double damp, int iters, double tol, int* done, double* tmp, double* sc, bool sync) {
  hipStream_t s = stream();
  fwd_prepare<TT>(P, s);
  fwd_init<TT>(P, s, tmp, sc);
  int it = 0;
  double res = 0.0;
  for (it = 0; it < iters; ++it) {
    fwd_iter<TT>(P, damp, sc + 2 * it, sc + 2 * (it + 1), s, tol > 0.0 || it + 1 == iters);
    if (tol > 0.0) {
      double h[2];
      GRE_HIP(hipMemcpyAsync(h, sc + 2 * (it + 1), sizeof(h), hipMemcpyDeviceToHost, s));
      GRE_HIP(hipStreamSynchronize(s));
      res = h[1];
      if (res < tol) { ++it; break; }
    }
  }
  fwd_done(P);
  if (done) *done = it;
  if (!sync && tol <= 0.0) return 0.0;
  if (tol <= 0.0 && iters > 0) {
    double h[2];
    GRE_HIP(hipMemcpyAsync(h, sc + 2 * iters, sizeof(h), hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
    res = h[1];
  }
  GRE_HIP(hipStreamSynchronize(s));
  return res;
}

// Applies schedule code sc (GRE_SCHED_*) and tile size to the handle.
void apply_schedule(gre_pr* P, int sc, int ipt) {
  P->schedule = sc;
  P->ipt = ipt;
  P->nt = sc == GRE_SCHED_TILE_NT || sc == GRE_SCHED_COMPACT_NT || sc == GRE_SCHED_PART_NT;
  P->part = sc == GRE_SCHED_PART || sc == GRE_SCHED_PART_NT;
  P->compact = sc == GRE_SCHED_COMPACT || sc == GRE_SCHED_COMPACT_NT || P->part;
  P->bin = sc == GRE_SCHED_BIN;
}

// Autotuner: times every (schedule, tile size) candidate on two forward
// iterations after a warm-up one and keeps the fastest.  Candidates: tile
// sizes of 4/8/16 entries per thread, with and without the non-temporal id
// stream, and for the single vector (weighted or not) the compact and the
// partitioned schedules; for the unweighted FP32 single vector also the
// binned pull.
template <typename TT>
void autotune_forward(gre_pr* P, double damp) {
  hipStream_t s = stream();
  const int ipts[3] = {4, 8, 16};
  std::vector<int> scheds = {GRE_SCHED_TILE, GRE_SCHED_TILE_NT};
  if (P->k == 1) {
    scheds.push_back(GRE_SCHED_COMPACT);
    scheds.push_back(GRE_SCHED_COMPACT_NT);
  }
  const bool can_part = P->k == 1 && P->g->nE > 0 && P->g->nE < 0xffffffffULL;
  int best_s = GRE_SCHED_TILE, best_i = 8, best_r = P->ranges;
  uint32_t best_h = P->shared;
  float best_ms = 1e30f;
  hipEvent_t a, b;
  GRE_HIP(hipEventCreate(&a));
  GRE_HIP(hipEventCreate(&b));
  ensure_scal(P, 4);
  // candidates' partitions stay cached while tuning; cleared on every exit
  struct Tuning {
    gre_pr* P;
    explicit Tuning(gre_pr* p) : P(p) { P->tuning = true; }
    ~Tuning() { P->tuning = false; }
  } tuning_scope(P);
  // one candidate: a warm iteration, then `timed` timed ones; returns their ms
  auto time_it = [&](int sc, int ipt, bool prt, const std::pair<int, uint32_t>& h, int timed) {
    apply_schedule(P, sc, ipt);
    if (prt) { P->ranges = h.first; P->shared = h.second; }
    ensure_plan(P, true, ipt);
    double* scal = P->scal.as<double>();
    fwd_init<TT>(P, s, static_cast<double*>(scratch(3, fwd_init_tmp_bytes(P->g->nV))), scal);
    fwd_iter<TT>(P, damp, scal, scal + 2, s);  // warm
    GRE_HIP(hipEventRecord(a, s));
    for (int i = 0; i < timed; ++i) fwd_iter<TT>(P, damp, scal + 2 + 2 * (i & 1), scal + 4 - 2 * (i & 1), s);
    GRE_HIP(hipEventRecord(b, s));
    GRE_HIP(hipEventSynchronize(b));
    float ms = 0.f;
    GRE_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
  };
  auto trial = [&](int sc, int ipt, bool prt, const std::pair<int, uint32_t>& h) {
    const float ms = time_it(sc, ipt, prt, h, 2);
    if (ms < best_ms) {
      best_ms = ms;
      best_s = sc;
      best_i = ipt;
      if (prt) { best_r = h.first; best_h = h.second; }
    }
    return ms;
  };
  for (int sc : scheds)
    for (int ipt : ipts) trial(sc, ipt, false, {P->ranges, P->shared});
  // the binned pull (FP32 single vector, unweighted): one layout, no tile size
  std::shared_ptr<XPart> bin_xp;
  {
    graph_compact_targets(P->g);  // the vertices with in-edges, which the bins hold
    if (!bin_refusal(P) && (uint64_t)P->g->ntgt <= (uint64_t)kBinMaxBins * kBinRows) {
      trial(GRE_SCHED_BIN, 8, false, {P->ranges, P->shared});
      bin_xp = P->xp;
      P->xp.reset();
    }
  }
  // partitioned schedules try (tail ranges, shared sources) pairs, one
  // partition at a time; the two best partitions (each with its best
  // schedule) stay cached and are timed again against each other at the end,
  // alternately, since a few percent separate them and two iterations right
  // after a partition build do not always rank them right
  struct PartCand {
    float ms = 1e30f;
    int sc = 0, ipt = 0;
    std::pair<int, uint32_t> h{0, 0};
    std::shared_ptr<XPart> xp;
  } top[2];
  if (can_part) {
    const std::pair<int, uint32_t> vars[5] = {{256, 1u << 12}, {128, 1u << 12}, {256, 0u}, {128, 1u << 17}, {64, 1u << 18}};
    for (const auto& h : vars) {
      PartCand c;
      c.h = h;
      for (int sc : {GRE_SCHED_PART, GRE_SCHED_PART_NT})
        for (int ipt : {4, 8}) {
          const float ms = trial(sc, ipt, true, h);
          if (ms < c.ms) { c.ms = ms; c.sc = sc; c.ipt = ipt; }
        }
      c.xp = P->xp;
      if (c.ms < top[0].ms) { top[1] = std::move(top[0]); top[0] = std::move(c); }
      else if (c.ms < top[1].ms) top[1] = std::move(c);
      P->xp.reset();
      P->wpart_xp.reset();
      P->wpart_ver = ~0ULL;
      P->plan_ipt_x = 0;
      prune_xparts(P->g);
    }
    const bool part_won = best_s == GRE_SCHED_PART || best_s == GRE_SCHED_PART_NT;
    if (part_won && top[0].xp && top[1].xp) {
      float tot[2] = {0.f, 0.f};
      for (int round = 0; round < 3; ++round)
        for (int k = 0; k < 2; ++k) {
          P->xp = top[k].xp;
          P->wpart_xp.reset();
          P->wpart_ver = ~0ULL;
          P->plan_ipt_x = 0;
          tot[k] += time_it(top[k].sc, top[k].ipt, true, top[k].h, 4);
        }
      if (tot[1] < tot[0]) std::swap(top[0], top[1]);
    }
    if (part_won && top[0].xp) {
      best_s = top[0].sc;
      best_i = top[0].ipt;
    }
    // the best partition's (ranges, shared) also serve a binned forward's adjoint
    if (top[0].xp) {
      best_r = top[0].h.first;
      best_h = top[0].h.second;
    }
  }
  std::shared_ptr<XPart> best_xp =
      (best_s == GRE_SCHED_PART || best_s == GRE_SCHED_PART_NT) ? top[0].xp
      : best_s == GRE_SCHED_BIN                                 ? bin_xp
                                                                : std::shared_ptr<XPart>();
  bin_xp.reset();
  top[1].xp.reset();
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  apply_schedule(P, best_s, best_i);
  P->ranges = best_r;
  P->shared = best_h;
  // the handle still holds the last candidate's partition: take the winner's
  // from the cache, then drop the candidates that lost (unless another handle
  // holds them)
  P->tuning = false;
  P->xp.reset();
  P->wpart_xp.reset();
  P->wpart_ver = ~0ULL;
  P->plan_ipt_x = 0;
  if (P->part || P->bin) ensure_plan(P, true, best_i);
  best_xp.reset();
  prune_xparts(P->g);
}

// the binned adjoint's layout, buffers, out-degrees and weights in its order,
// and the bounds of its fixed-point sums
void ensure_bin_adj_plan(gre_pr* P, int iters) {
  hipStream_t s = stream();
  gre_graph* g = P->g;
  const uint32_t nV = g->nV;
  if (!P->xpb) {
    P->xpb = graph_binpart(graph_reversed(g));
    P->wbinb_xp.reset();
    P->wbinb_ver = ~0ULL;
    P->maxdeg_b = 0.0;
  }
  const XPart& xp = *P->xpb;
  if (P->bbufb.for_xp != &xp) P->bbufb.make(xp, nV, s);
  const size_t xb = sizeof(float) * (size_t)nV;
  if (P->cb_r.bytes < xb) {
    P->cb_r.alloc(xb);
    P->pb_r.alloc(xb);
    P->degb_r.alloc(xb);
    P->wlb_r.alloc(xb);
  }
  if (P->mxb.bytes < 4 * ((size_t)iters + 1)) P->mxb.alloc(4 * ((size_t)iters + 1));
  if (P->maxdeg_b == 0.0) {
    hipLaunchKernelGGL(k_permute32<int32_t>, dim3(gcap(nV)), dim3(kB), 0, s, g->outdeg.as<int32_t>(),
                       xp.rowinv.as<uint32_t>(), (uint64_t)nV, P->degb_r.as<int32_t>());
    GRE_HIP(hipGetLastError());
    uint32_t* m = static_cast<uint32_t*>(scratch(4, 16));
    bin_zero_u32(m, 1, s);
    bin_maxu(g->outdeg.as<uint32_t>(), nV, m, s);
    uint32_t h = 0;
    GRE_HIP(hipMemcpyAsync(&h, m, 4, hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
    P->maxdeg_b = std::max(1.0, (double)h);
  }
  if (P->weighted && (P->wbinb_xp != P->xpb || P->wbinb_ver != P->wver)) {
    // the normalised weights of the reversed graph's CSR order are the CSC's
    const uint64_t Ew = xp.bin.Ew;
    P->wbinb.alloc(sizeof(float) * Ew);
    hipLaunchKernelGGL(k_permute32<float>, dim3(gcap(Ew)), dim3(kB), 0, s, P->wcsc.as<float>(),
                       xp.bin.wpos.as<uint32_t>(), Ew, P->wbinb.as<float>());
    GRE_HIP(hipGetLastError());
    uint32_t* m = static_cast<uint32_t*>(scratch(4, 16));
    bin_zero_u32(m, 1, s);
    bin_absmax(P->wcsc.as<float>(), P->g->nE, m, s);
    uint32_t h = 0;
    GRE_HIP(hipMemcpyAsync(&h, m, 4, hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
    float w = 0.f;
    std::memcpy(&w, &h, 4);
    P->wmax_b = (double)w;
    P->wbinb_xp = P->xpb;
    P->wbinb_ver = P->wver;
  }
}

// The single-vector FP32 adjoint over the binned layout of the reversed graph:
// lambda, c, p and (W lambda) kept in its row order, turned back into vertex
// order at the end.  Each iteration's fixed-point scale follows max |lambda|
// of the iteration before (measured on the device).
double run_adjoint_bin(gre_pr* P, double damp, int iters, double tol, int* done) {
  hipStream_t s = stream();
  gre_graph* g = P->g;
  const uint32_t nV = g->nV;
  const unsigned gb = gcap(nV);
  ensure_bin_adj_plan(P, iters);
  ensure_scal(P, iters);
  double* tmp = static_cast<double*>(scratch(3, sizeof(double) * 2 * gb));
  const XPart& xp = *P->xpb;
  const uint32_t* rinv = xp.rowinv.as<uint32_t>();
  const uint32_t* rperm = xp.rowperm.as<uint32_t>();
  hipLaunchKernelGGL(k_permute32<float>, dim3(gb), dim3(kB), 0, s, P->c.as<float>(), rinv, (uint64_t)nV,
                     P->cb_r.as<float>());
  if (P->has_p)
    hipLaunchKernelGGL(k_permute32<float>, dim3(gb), dim3(kB), 0, s, P->p.as<float>(), rinv, (uint64_t)nV,
                       P->pb_r.as<float>());
  uint32_t* mx = P->mxb.as<uint32_t>();
  bin_zero_u32(mx, (uint32_t)iters + 1, s);
  double* sc = P->scal.as<double>();
  P->lcur = 0;
  hipLaunchKernelGGL(k_adj_init<float>, dim3(gb), dim3(kB), 0, s, nV, P->cb_r.as<float>(),
                     P->has_p ? P->pb_r.as<float>() : nullptr, P->lam[0].as<float>(), 1.0 / (double)nV, tmp,
                     (const uint32_t*)nullptr, (float*)nullptr);
  hipLaunchKernelGGL(k_reduce_pairs, dim3(1), dim3(1024), 0, s, tmp, (uint64_t)gb, tmp, (uint64_t)0, sc);
  GRE_HIP(hipGetLastError());
  bin_absmax(P->lam[0].as<float>(), nV, mx, s);
  const double F = P->maxdeg_b * (P->weighted ? P->wmax_b : 1.0);
  int it = 0;
  double res = 0.0;
  for (it = 0; it < iters; ++it) {
    BinAdjIter a;
    a.wbin = P->weighted ? P->wbinb.as<float>() : nullptr;
    a.lin = P->lam[P->lcur].as<float>();
    a.lout = P->lam[P->lcur ^ 1].as<float>();
    a.wl = P->wlb_r.as<float>();
    a.c_r = P->cb_r.as<float>();
    a.p_r = P->has_p ? P->pb_r.as<float>() : nullptr;
    a.deg_r = P->degb_r.as<int32_t>();
    a.scal_in = sc + 2 * it;
    a.scal_out = sc + 2 * (it + 1);
    a.mx_in = mx + it;
    a.mx_out = mx + it + 1;
    a.F = F;
    a.damp = damp;
    a.invn = 1.0 / (double)nV;
    bin_adjoint_iter(xp, P->bbufb, a, nV, s);
    P->lcur ^= 1;
    if (tol > 0.0) {
      double h[2];
      GRE_HIP(hipMemcpyAsync(h, sc + 2 * (it + 1), sizeof(h), hipMemcpyDeviceToHost, s));
      GRE_HIP(hipStreamSynchronize(s));
      res = h[1];
      if (res < tol) { ++it; break; }
    }
  }
  if (tol <= 0.0 && iters > 0) {
    double h[2];
    GRE_HIP(hipMemcpyAsync(h, sc + 2 * iters, sizeof(h), hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
    res = h[1];
  }
  // back to vertex order
  hipLaunchKernelGGL(k_permute32<float>, dim3(gb), dim3(kB), 0, s, P->lam[P->lcur].as<float>(), rperm,
                     (uint64_t)nV, P->lam[P->lcur ^ 1].as<float>());
  P->lcur ^= 1;
  if (it > 0)
    hipLaunchKernelGGL(k_permute32<float>, dim3(gb), dim3(kB), 0, s, P->wlb_r.as<float>(), rperm, (uint64_t)nV,
                       P->wl.as<float>());
  GRE_HIP(hipGetLastError());
  GRE_HIP(hipStreamSynchronize(s));
  if (done) *done = it;
  return res;
}

template <typename TT>
double run_adjoint(gre_pr* P, double damp, int iters, double tol, int* done) {
  if constexpr (std::is_same<TT, float>::value) {
    if (adj_bin(P)) return run_adjoint_bin(P, damp, iters, tol, done);
  }
  hipStream_t s = stream();
  const uint32_t nV = P->g->nV;
  const unsigned gb = gcap(nV);
  double* tmp = static_cast<double*>(scratch(3, sizeof(double) * 2 * gb));
  ensure_plan(P, false, P->ipt);
  ensure_scal(P, iters);
  const bool cpt = adj_compact(P);
  if (cpt && P->lc[0].bytes < tsize(P->dtype) * (size_t)nV) {
    P->lc[0].alloc(tsize(P->dtype) * (size_t)nV);
    P->lc[1].alloc(tsize(P->dtype) * (size_t)nV);
  }
  const uint32_t* tid = cpt ? P->g->did.as<uint32_t>() : nullptr;
  double* sc = P->scal.as<double>();
  P->lcur = 0;
  hipLaunchKernelGGL(k_adj_init<TT>, dim3(gb), dim3(kB), 0, s, nV, P->c.as<TT>(), P->has_p ? P->p.as<TT>() : nullptr,
                     P->lam[0].as<TT>(), 1.0 / (double)nV, tmp, tid, cpt ? P->lc[0].as<TT>() : nullptr);
  hipLaunchKernelGGL(k_reduce_pairs, dim3(1), dim3(1024), 0, s, tmp, (uint64_t)gb, tmp, (uint64_t)0, sc);
  GRE_HIP(hipGetLastError());
  int it = 0;
  double res = 0.0;
  for (it = 0; it < iters; ++it) {
    ModeAdj<TT> m;
    m.lin = P->lam[P->lcur].as<TT>();
    m.lout = P->lam[P->lcur ^ 1].as<TT>();
    m.gin = cpt ? P->lc[P->lcur].as<TT>() : m.lin;
    m.lc_out = cpt ? P->lc[P->lcur ^ 1].as<TT>() : nullptr;
    m.tid = tid;
    m.w = P->weighted ? P->wcsr.as<TT>() : nullptr;
    m.unw = !P->weighted;
    m.c = P->c.as<TT>();
    m.p = P->has_p ? P->p.as<TT>() : nullptr;
    m.outdeg = P->g->outdeg.as<int32_t>();
    m.wl = P->wl.as<TT>();
    m.scal_in = sc + 2 * it;
    m.damp = damp;
    m.invn = 1.0 / (double)nV;
    if (adj_part(P)) {
      // segment sums over XCD-local ranges of compact targets, then each row's segments
      if (P->weighted)
        pull_pass(P->plan_xa, ModeSeg<TT, true>{m.gin, P->asum.as<TT>(), P->wparta.as<TT>()}, nullptr, s);
      else
        pull_pass(P->plan_xa, ModeSeg<TT, false>{m.gin, P->asum.as<TT>(), nullptr}, nullptr, s);
      m.gin = P->asum.as<TT>();
      m.w = nullptr;
      adj_pull(P->plan_pxa, m, sc + 2 * (it + 1), s);
    } else {
      adj_pull(P->plan_csr, m, sc + 2 * (it + 1), s);
    }
    P->lcur ^= 1;
    if (tol > 0.0) {
      double h[2];
      GRE_HIP(hipMemcpyAsync(h, sc + 2 * (it + 1), sizeof(h), hipMemcpyDeviceToHost, s));
      GRE_HIP(hipStreamSynchronize(s));
      res = h[1];
      if (res < tol) { ++it; break; }
    }
  }
  if (tol <= 0.0 && iters > 0) {
    double h[2];
    GRE_HIP(hipMemcpyAsync(h, sc + 2 * iters, sizeof(h), hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
    res = h[1];
  }
  GRE_HIP(hipStreamSynchronize(s));
  if (done) *done = it;
  return res;
}

template <typename TT>
void gradients(gre_pr* P, double damp, float* gW_host, float* graw_host, void* gp_host) {
  hipStream_t s = stream();
  gre_graph* g = P->g;
  ranks_to_vertices(P);
  const TT* x = P->x[P->xcur].as<TT>();
  const TT* lam = P->lam[P->lcur].as<TT>();
  // dangling mass of the final forward, recomputed in fixed order
  const unsigned gb = gcap(g->nV);
  double* tmp = static_cast<double*>(scratch(3, sizeof(double) * 2 * gb));
  double* dsc = static_cast<double*>(scratch(4, sizeof(double) * 2));
  DevBuf dummy;
  dummy.alloc(sizeof(TT) * (size_t)g->nV);
  hipLaunchKernelGGL(k_fwd_init<TT>, dim3(gb), dim3(kB), 0, s, g->nV, x, g->outdeg.as<int32_t>(), dummy.as<TT>(),
                     (TT*)nullptr, 0.0, tmp, (const uint32_t*)nullptr, false);
  hipLaunchKernelGGL(k_reduce_pairs, dim3(1), dim3(1024), 0, s, tmp, (uint64_t)gb, tmp, (uint64_t)0, dsc);
  if (gp_host) {
    hipLaunchKernelGGL(k_grad_p<TT>, dim3(gb), dim3(kB), 0, s, g->nV, lam, dsc, damp, dummy.as<TT>());
    GRE_HIP(hipMemcpyAsync(gp_host, dummy.p, sizeof(TT) * g->nV, hipMemcpyDeviceToHost, s));
  }
  if (gW_host || graw_host) {
    const uint32_t te = 2048;
    uint64_t nt = 0;
    const uint32_t* trb = graph_tiles(g, false, te, &nt);
    DevBuf dgW, dgraw;
    dgW.alloc(4 * (size_t)g->nE);
    if (graw_host) dgraw.alloc(4 * (size_t)g->nE);
    if (g->nE > 0) {
      ProfScope ps("edge_grad");
      hipLaunchKernelGGL(k_edge_grad<TT>, dim3((unsigned)nt), dim3(256), 0, s, g->off.as<uint64_t>(),
                         g->idx.as<uint32_t>(), trb, g->nE, g->nV, nt, te,
                         P->weighted ? P->wcsr.p : nullptr, P->weighted ? P->rowsum.as<double>() : nullptr,
                         x, lam, P->wl.as<TT>(), g->outdeg.as<int32_t>(), damp, dgW.as<float>(),
                         graw_host ? dgraw.as<float>() : nullptr);
    }
    GRE_HIP(hipGetLastError());
    if (gW_host && g->nE) GRE_HIP(hipMemcpyAsync(gW_host, dgW.p, 4 * g->nE, hipMemcpyDeviceToHost, s));
    if (graw_host && g->nE) GRE_HIP(hipMemcpyAsync(graw_host, dgraw.p, 4 * g->nE, hipMemcpyDeviceToHost, s));
  }
  GRE_HIP(hipStreamSynchronize(s));
}

void alloc_state(gre_pr* P) {
  const size_t n = (size_t)P->g->nV * P->k;
  const size_t ts = tsize(P->dtype);
  P->x[0].alloc(ts * n);
  if (P->weighted || P->k > 1) P->x[1].alloc(ts * n);
  if (P->k == 1) { P->contrib[0].alloc(ts * n); P->contrib[1].alloc(ts * n); }
}

void alloc_adjoint(gre_pr* P) {
  if (P->c.p) return;
  const size_t bytes = tsize(P->dtype) * (size_t)P->g->nV * P->k;
  P->c.alloc(bytes);
  P->lam[0].alloc(bytes);
  P->lam[1].alloc(bytes);
  P->wl.alloc(bytes);
}

// Transposed power iteration from the seed vector already in P->c.
void solve_adjoint(gre_pr* P, double damping, int iters, double tol, int* iters_done, double* residual) {
  double r;
  if (P->k > 1) {
    r = batched_adjoint(P, damping, iters, tol, iters_done);
  } else {
    if (P->ipt == 0) P->ipt = 8;
    r = P->dtype == GRE_F64 ? run_adjoint<double>(P, damping, iters, tol, iters_done)
                            : run_adjoint<float>(P, damping, iters, tol, iters_done);
  }
  if (residual) *residual = r;
  P->have_adj = true;
}

// The address of `scratch` on the current device, looked up once per device
// (each GPU holds its own copy of the array at its own address).
double* noalloc_scratch() {
  static std::mutex mu;
  static double* ptr[64] = {};
  int dev = 0;
  GRE_HIP(hipGetDevice(&dev));
  GRE_CHECK(dev >= 0 && dev < 64, GRE_DEV_FAILURE, "device index out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    GRE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
    GRE_CHECK(p, GRE_DEV_FAILURE, "no device address for the allocation-free scratch");
    ptr[dev] = static_cast<double*>(p);
  }
  return ptr[dev];
}

// True when the forward's pull plan for the handle's current schedule, tile
// size, partition and weights is built, i.e. ensure_plan(P, true, P->ipt)
// would allocate nothing and launch nothing.
bool fwd_plan_ready(const gre_pr* P) {
  if (P->ipt == 0) return false;                      // schedule not tuned yet
  if (use_bin(P)) {
    if (bin_refusal(P) || !P->xp || !P->xp->binned || P->bbuf.for_xp != P->xp.get()) return false;
    if (P->weighted && (P->wpart_xp != P->xp || P->wpart_ver != P->wver)) return false;
    return P->x[1].bytes >= sizeof(float) * (size_t)P->g->nV && P->p_r.bytes >= sizeof(float) * (size_t)P->g->nV;
  }
  if (use_compact(P) && !P->g->cid.p) return false;
  if (use_part(P)) {
    if (!P->xp || P->xp->R != P->ranges || P->xp->H != std::min(P->shared, P->g->nsrc)) return false;
    if (P->plan_ipt_x != P->ipt) return false;
    if (P->xsum.bytes < tsize(P->dtype) * (size_t)P->xp->nseg) return false;
    if (P->x[1].bytes < tsize(P->dtype) * (size_t)P->g->nV || P->p_r.bytes < tsize(P->dtype) * (size_t)P->g->nV)
      return false;
    if (P->weighted && (P->wpart_xp != P->xp || P->wpart_ver != P->wver)) return false;
    return true;
  }
  return P->plan_ipt_csc == P->ipt;
}

}  // namespace
}  // namespace gre

using namespace gre;

extern "C" {

int gre_pr_create(gre_pr** out, gre_graph* g, int dtype, int k, int weighted) {
  return guard([&] {
    GRE_CHECK(out && g, GRE_DEV_NULL, "null argument");
    GRE_CHECK(dtype == GRE_F32 || dtype == GRE_F64, GRE_DEV_NULL, "dtype must be GRE_F32 or GRE_F64");
    GRE_CHECK(k >= 1 && k <= kMaxBatch, GRE_DEV_NULL, "k must be in [1, 64]");
    GRE_CHECK(k == 1 || weighted, GRE_DEV_NULL, "batched vectors (k > 1) need weighted = 1");
    auto* P = new gre_pr();
    try {
      P->g = g;
      P->dtype = dtype;
      P->k = k;
      P->weighted = weighted ? 1 : 0;
      alloc_state(P);
      if (P->weighted) {
        // unit raw weights until set: row-normalised 1/outdeg
        const int rc = gre_pr_set_edge_weights(P, nullptr, 0);
        GRE_CHECK(rc == 0, rc, gre_last_error());
      }
    } catch (...) {
      delete P;
      throw;
    }
    *out = P;
  });
}

int gre_pr_destroy(gre_pr* P) {
  return guard([&] {
    if (!P) return;
    GRE_HIP(hipStreamSynchronize(stream()));
    // the graph keeps this handle's partitions cached for the next handle and
    // drops the others that nothing holds any more
    gre_graph* g = P->g;
    std::shared_ptr<XPart> keep_f = P->xp, keep_a = P->xpa;
    delete P;
    prune_xparts(g);
  });
}

int gre_pr_set_graphs(gre_pr* P, int on) {
  return guard([&] {
    GRE_CHECK(P, GRE_DEV_NULL, "null handle");
    P->graphs = on != 0;
  });
}

int gre_pr_set_schedule(gre_pr* P, int schedule, int tile_items) {
  return guard([&] {
    GRE_CHECK(P, GRE_DEV_NULL, "null handle");
    GRE_CHECK(schedule >= GRE_SCHED_AUTO && schedule <= GRE_SCHED_BIN, GRE_DEV_NULL,
              "schedule must be one of GRE_SCHED_*");
    GRE_CHECK(schedule != GRE_SCHED_BIN || P->k == 1, GRE_DEV_NULL, "GRE_SCHED_BIN runs single-vector handles only");
    GRE_CHECK(P->k > 1 ? (schedule == GRE_SCHED_AUTO || schedule == GRE_SCHED_TILE || schedule == GRE_SCHED_ROWS)
                       : schedule != GRE_SCHED_ROWS,
              GRE_DEV_NULL, "batched handles take GRE_SCHED_AUTO, _TILE or _ROWS; GRE_SCHED_ROWS needs k > 1");
    GRE_CHECK(tile_items == 0 || tile_items == 4 || tile_items == 8 || tile_items == 16, GRE_DEV_NULL,
              "tile items per thread must be 4, 8 or 16");
    if (schedule == GRE_SCHED_AUTO) apply_schedule(P, GRE_SCHED_AUTO, 0);
    else apply_schedule(P, schedule, tile_items ? tile_items : 8);
  });
}

int gre_pr_get_schedule(gre_pr* P, int* schedule, int* tile_items) {
  return guard([&] {
    GRE_CHECK(P && schedule && tile_items, GRE_DEV_NULL, "null argument");
    *schedule = P->schedule;
    *tile_items = P->ipt;
    // batched handles under AUTO: the schedule the forward timing picked
    if (P->k > 1 && P->schedule == GRE_SCHED_AUTO && P->bstate.sched[0]) *schedule = P->bstate.sched[0];
  });
}

int gre_pr_set_partition(gre_pr* P, int ranges, uint32_t shared_sources) {
  return guard([&] {
    GRE_CHECK(P, GRE_DEV_NULL, "null handle");
    GRE_CHECK(ranges == 8 || ranges == 16 || ranges == 32 || ranges == 64 || ranges == 128 || ranges == 256,
              GRE_DEV_NULL, "source ranges must be 8, 16, 32, 64, 128 or 256");
    P->ranges = ranges;
    P->shared = shared_sources;
  });
}

int gre_pr_get_partition(gre_pr* P, int* ranges, uint32_t* shared_sources) {
  return guard([&] {
    GRE_CHECK(P && ranges && shared_sources, GRE_DEV_NULL, "null argument");
    *ranges = P->ranges;
    *shared_sources = P->shared;
  });
}

int gre_pr_set_edge_weights(gre_pr* P, const float* w, uint64_t seed) {
  return guard([&] {
    GRE_CHECK(P, GRE_DEV_NULL, "null handle");
    GRE_CHECK(P->weighted, GRE_DEV_NULL, "handle was created unweighted");
    gre_graph* g = P->g;
    hipStream_t s = stream();
    const uint64_t nE = g->nE;
    // raw weights are FP32 (the input); normalised weights use the compute type
    const size_t ts = tsize(P->dtype);
    P->wraw.alloc(4 * nE);
    P->wcsr.alloc(ts * nE);
    P->wcsc.alloc(ts * nE);
    P->rowsum.alloc(8 * (size_t)g->nV);
    if (nE > 0) {
      if (w) {
        GRE_HIP(hipMemcpyAsync(P->wraw.p, w, 4 * nE, hipMemcpyHostToDevice, s));
      } else if (seed == 0) {
        std::vector<float> ones(nE, 1.0f);
        GRE_HIP(hipMemcpyAsync(P->wraw.p, ones.data(), 4 * nE, hipMemcpyHostToDevice, s));
        GRE_HIP(hipStreamSynchronize(s));
      } else {
        hipLaunchKernelGGL(k_lognormal, dim3(gcap(nE)), dim3(kB), 0, s, g->off.as<uint64_t>(), g->idx.as<uint32_t>(),
                           g->nV, nE, seed, P->wraw.as<float>());
      }
    }
    // row sums in FP64 (pull over CSR with gather = weight)
    {
      PullPlan pl;
      PullBuffers pb;
      pb.make(pl, g, false, 8, sizeof(double));
      ModeRowSum m{P->wraw.as<float>(), P->rowsum.as<double>()};
      double* sc = static_cast<double*>(gre::scratch(4, 2 * sizeof(double)));
      pull_pass(pl, m, sc, s);
      GRE_HIP(hipStreamSynchronize(s));
    }
    if (nE > 0) {
      const uint64_t* ceid = graph_ceid(g);
      if (P->dtype == GRE_F64) {
        hipLaunchKernelGGL(k_normalise<double>, dim3(gcap(nE)), dim3(kB), 0, s, g->off.as<uint64_t>(), g->nV, nE,
                           P->wraw.as<float>(), P->rowsum.as<double>(), P->wcsr.as<double>());
        hipLaunchKernelGGL(k_permute<double>, dim3(gcap(nE)), dim3(kB), 0, s, P->wcsr.as<double>(), ceid, nE,
                           P->wcsc.as<double>());
      } else {
        hipLaunchKernelGGL(k_normalise<float>, dim3(gcap(nE)), dim3(kB), 0, s, g->off.as<uint64_t>(), g->nV, nE,
                           P->wraw.as<float>(), P->rowsum.as<double>(), P->wcsr.as<float>());
        hipLaunchKernelGGL(k_permute<float>, dim3(gcap(nE)), dim3(kB), 0, s, P->wcsr.as<float>(), ceid, nE,
                           P->wcsc.as<float>());
      }
      GRE_HIP(hipGetLastError());
    }
    ++P->wver;
    GRE_HIP(hipStreamSynchronize(s));
  });
}

int gre_pr_get_edge_weights(gre_pr* P, float* w_raw, float* w_norm) {
  return guard([&] {
    GRE_CHECK(P, GRE_DEV_NULL, "null handle");
    GRE_CHECK(P->weighted, GRE_DEV_NULL, "handle was created unweighted");
    hipStream_t s = stream();
    const uint64_t nE = P->g->nE;
    if (w_raw && nE) GRE_HIP(hipMemcpyAsync(w_raw, P->wraw.p, 4 * nE, hipMemcpyDeviceToHost, s));
    if (w_norm && nE) {
      if (P->dtype == GRE_F64) {
        std::vector<double> h(nE);
        GRE_HIP(hipMemcpyAsync(h.data(), P->wcsr.p, 8 * nE, hipMemcpyDeviceToHost, s));
        GRE_HIP(hipStreamSynchronize(s));
        for (uint64_t e = 0; e < nE; ++e) w_norm[e] = (float)h[e];
      } else {
        GRE_HIP(hipMemcpyAsync(w_norm, P->wcsr.p, 4 * nE, hipMemcpyDeviceToHost, s));
      }
    }
    GRE_HIP(hipStreamSynchronize(s));
  });
}

int gre_pr_set_personalization(gre_pr* P, const void* pvec) {
  return guard([&] {
    GRE_CHECK(P, GRE_DEV_NULL, "null handle");
    if (!pvec) { P->has_p = false; P->p_l1 = 1.0; P->p.release(); return; }
    const size_t bytes = tsize(P->dtype) * (size_t)P->g->nV * P->k;
    P->p.alloc(bytes);
    GRE_HIP(hipMemcpyAsync(P->p.p, pvec, bytes, hipMemcpyHostToDevice, stream()));
    // the largest column L1 norm (the binned schedule's fixed-point sums are bounded by it)
    double l1 = 0.0;
    const size_t n = (size_t)P->g->nV;
    for (int j = 0; j < P->k; ++j) {
      double c = 0.0;
      for (size_t v = 0; v < n; ++v)
        c += std::fabs(P->dtype == GRE_F64 ? static_cast<const double*>(pvec)[v * P->k + j]
                                           : (double)static_cast<const float*>(pvec)[v * P->k + j]);
      l1 = std::max(l1, c);
    }
    P->p_l1 = l1;
    GRE_HIP(hipStreamSynchronize(stream()));
    P->has_p = true;
    ++P->pver;
  });
}

int gre_pr_run(gre_pr* P, double damping, int iters, double tol, int* iters_done, double* residual) {
  return guard([&] {
    GRE_CHECK(P, GRE_DEV_NULL, "null handle");
    GRE_CHECK(iters >= 0, GRE_DEV_NULL, "iters must be >= 0");
    GRE_CHECK(damping >= 0.0 && damping < 1.0, GRE_DEV_NULL, "damping must be in [0,1)");
    if (P->k > 1) {
      GRE_CHECK(P->has_p, GRE_DEV_NULL, "batched runs need a personalisation matrix");
      double r = batched_forward(P, damping, iters, tol, iters_done);
      if (residual) *residual = r;
      P->have_fwd = true;
      return;
    }
    if (P->ipt == 0) {
      if (P->dtype == GRE_F64) autotune_forward<double>(P, damping);
      else autotune_forward<float>(P, damping);
    }
    ensure_plan(P, true, P->ipt);
    ensure_scal(P, iters);
    const double r = P->dtype == GRE_F64 ? run_forward<double>(P, damping, iters, tol, iters_done)
                                         : run_forward<float>(P, damping, iters, tol, iters_done);
    if (residual) *residual = r;
    P->have_fwd = true;
  });
}

int gre_pr_run_noalloc(gre_pr* P, double damping, int iters, double tol, int* iters_done, double* residual) {
  return guard([&] {
    GRE_CHECK(P, GRE_DEV_NULL, "null handle");
    GRE_CHECK(iters >= 0, GRE_DEV_NULL, "iters must be >= 0");
    GRE_CHECK(damping >= 0.0 && damping < 1.0, GRE_DEV_NULL, "damping must be in [0,1)");
    GRE_CHECK(P->k == 1, GRE_DEV_NOT_IMPLEMENTED, "gre_pr_run_noalloc runs single-vector handles only (k = 1)");
    GRE_CHECK(iters <= kNoAllocIters, GRE_DEV_NOBUFFER,
              "gre_pr_run_noalloc runs at most 4096 iterations (its scratch holds 4097 reduction pairs)");
    GRE_CHECK(!prof_on(), GRE_DEV_FAILURE, "gre_pr_run_noalloc does not run with kernel profiling on (it records events)");
    GRE_CHECK(fwd_plan_ready(P), GRE_DEV_NOBUFFER,
              "the handle's pull plan for its schedule is not built yet: one gre_pr_run on it (which allocates) "
              "prepares it");
    GRE_CHECK(fwd_init_tmp_bytes(P->g->nV) <= sizeof(double) * kNoAllocTmp, GRE_DEV_FAILURE, "init partials exceed scratch");
    double* base = noalloc_scratch();
    ensure_plan(P, true, P->ipt);   // ready: sets the plan's flags only
    const bool sync = residual != nullptr || tol > 0.0;
    const double r = P->dtype == GRE_F64
                         ? forward_eager<double>(P, damping, iters, tol, iters_done, base, base + kNoAllocTmp, sync)
                         : forward_eager<float>(P, damping, iters, tol, iters_done, base, base + kNoAllocTmp, sync);
    if (residual) *residual = r;
    P->have_fwd = true;
  });
}

int gre_pr_get_ranks(gre_pr* P, void* host_out) {
  return guard([&] {
    GRE_CHECK(P && host_out, GRE_DEV_NULL, "null argument");
    const size_t bytes = tsize(P->dtype) * (size_t)P->g->nV * P->k;
    ranks_to_vertices(P);
    GRE_HIP(hipMemcpyAsync(host_out, P->x[P->xcur].p, bytes, hipMemcpyDeviceToHost, stream()));
    GRE_HIP(hipStreamSynchronize(stream()));
  });
}

int gre_pr_adjoint(gre_pr* P, const uint32_t* targets, int64_t ntgt, double damping, int iters, double tol,
                   int* iters_done, double* residual) {
  return guard([&] {
    GRE_CHECK(P, GRE_DEV_NULL, "null handle");
    GRE_CHECK(P->have_fwd, GRE_DEV_NULL, "run the forward (gre_pr_run) first");
    GRE_CHECK(ntgt >= 0 && (ntgt == 0 || targets), GRE_DEV_NULL, "bad targets");
    for (int64_t i = 0; i < ntgt; ++i) GRE_CHECK(targets[i] < P->g->nV, GRE_DEV_NULL, "target id >= nV");
    hipStream_t s = stream();
    const size_t n = (size_t)P->g->nV * P->k;
    const size_t ts = tsize(P->dtype);
    alloc_adjoint(P);
    GRE_HIP(hipMemsetAsync(P->c.p, 0, ts * n, s));
    if (ntgt > 0) {
      DevBuf dt;
      dt.alloc(4 * (size_t)ntgt);
      GRE_HIP(hipMemcpyAsync(dt.p, targets, 4 * (size_t)ntgt, hipMemcpyHostToDevice, s));
      if (P->k == 1) {
        if (P->dtype == GRE_F64)
          hipLaunchKernelGGL(k_targets<double>, dim3(gcap(ntgt)), dim3(kB), 0, s, P->g->nV, dt.as<uint32_t>(), ntgt,
                             P->c.as<double>());
        else
          hipLaunchKernelGGL(k_targets<float>, dim3(gcap(ntgt)), dim3(kB), 0, s, P->g->nV, dt.as<uint32_t>(), ntgt,
                             P->c.as<float>());
      } else {
        batched_targets(P, dt.as<uint32_t>(), ntgt);
      }
      GRE_HIP(hipGetLastError());
      GRE_HIP(hipStreamSynchronize(s));
    }
    solve_adjoint(P, damping, iters, tol, iters_done, residual);
  });
}

int gre_pr_adjoint_cotangent(gre_pr* P, const void* cot, double damping, int iters, double tol, int* iters_done,
                             double* residual) {
  return guard([&] {
    GRE_CHECK(P && cot, GRE_DEV_NULL, "null argument");
    GRE_CHECK(P->have_fwd, GRE_DEV_NULL, "run the forward (gre_pr_run) first");
    hipStream_t s = stream();
    alloc_adjoint(P);
    GRE_HIP(hipMemcpyAsync(P->c.p, cot, tsize(P->dtype) * (size_t)P->g->nV * P->k, hipMemcpyHostToDevice, s));
    GRE_HIP(hipStreamSynchronize(s));
    solve_adjoint(P, damping, iters, tol, iters_done, residual);
  });
}

int gre_pr_get_adjoint(gre_pr* P, void* host_out) {
  return guard([&] {
    GRE_CHECK(P && host_out, GRE_DEV_NULL, "null argument");
    GRE_CHECK(P->have_adj, GRE_DEV_NULL, "run gre_pr_adjoint first");
    const size_t bytes = tsize(P->dtype) * (size_t)P->g->nV * P->k;
    GRE_HIP(hipMemcpyAsync(host_out, P->lam[P->lcur].p, bytes, hipMemcpyDeviceToHost, stream()));
    GRE_HIP(hipStreamSynchronize(stream()));
  });
}

int gre_pr_gradients(gre_pr* P, double damping, float* gW, float* graw, void* gp) {
  return guard([&] {
    GRE_CHECK(P, GRE_DEV_NULL, "null handle");
    GRE_CHECK(P->have_adj, GRE_DEV_NULL, "run gre_pr_adjoint first");
    if (P->k > 1) { batched_gradients(P, damping, gW, graw, gp); return; }
    if (P->dtype == GRE_F64) gradients<double>(P, damping, gW, graw, gp);
    else gradients<float>(P, damping, gW, graw, gp);
  });
}

}  // extern "C"
