// Graph construction on the device: R-MAT generation, CSR/CSC builds,
// SSSP weights, tile partitions for the pull kernels.
//
// GRE builds its local CSR by count -> scan -> fill in input order
// (include/Graph/LocalGraph.hpp:66-107) and its reverse CSR likewise with the
// input index kept in redgeID (:112-156).  Here both come from stable radix
// sorts, which give the same row contents and the same in-row order.
#include "graph.hpp"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

namespace gre {

namespace {

constexpr unsigned kBlock = 256;

// quadrant thresholds for (a,b,c,d) = (0.57,0.19,0.19,0.05) on 32-bit draws
constexpr uint64_t kTA = (57ULL << 32) / 100;
constexpr uint64_t kTAB = (76ULL << 32) / 100;
constexpr uint64_t kTABC = (95ULL << 32) / 100;

__device__ inline void rmat_edge(int scale, uint64_t seed, uint64_t i, uint32_t& u, uint32_t& v) {
  u = 0; v = 0;
  uint64_t h = 0;
  for (int l = 0; l < scale; ++l) {
    if ((l & 1) == 0) h = mix64(seed * GRE_K0 + i * GRE_K1 + (uint64_t)(l >> 1) * GRE_K2);
    const uint64_t r = (l & 1) ? (h & 0xffffffffULL) : (h >> 32);
    const uint32_t ub = r >= kTAB;
    const uint32_t vb = (r >= kTA && r < kTAB) || r >= kTABC;
    u = (u << 1) | ub;
    v = (v << 1) | vb;
  }
}

__global__ void k_rmat_keys(int scale, uint64_t seed, uint64_t ngen, int sym, int scramble, uint64_t* keys) {
  const uint64_t marker = ((1ULL << (2 * scale)) - 1);
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < ngen;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t u, v;
    rmat_edge(scale, seed, i, u, v);
    if (scramble) {
      u = rmat_scramble(scale, seed, u);
      v = rmat_scramble(scale, seed, v);
    }
    const uint64_t k1 = (u == v) ? marker : (((uint64_t)u << scale) | v);
    if (sym) {
      const uint64_t k2 = (u == v) ? marker : (((uint64_t)v << scale) | u);
      keys[2 * i] = k1;
      keys[2 * i + 1] = k2;
    } else {
      keys[i] = k1;
    }
  }
}

// For sorted packed keys (row << shift | col): idx[i] = col, and off[r] = first i of row r.
__global__ void k_keys_to_csr(const uint64_t* __restrict__ keys, uint64_t n, int shift, uint32_t nV,
                              uint64_t* __restrict__ off, uint32_t* __restrict__ idx) {
  const uint64_t mask = (1ULL << shift) - 1;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = keys[i];
    const uint32_t r = (uint32_t)(k >> shift);
    idx[i] = (uint32_t)(k & mask);
    const int64_t rprev = i == 0 ? -1 : (int64_t)(keys[i - 1] >> shift);
    for (int64_t q = rprev + 1; q <= (int64_t)r; ++q) off[q] = i;
    if (i == n - 1)
      for (uint64_t q = (uint64_t)r + 1; q <= nV; ++q) off[q] = n;
  }
}

// Same, for a sorted row-key array of uint32 (stable sorts of edge lists).
__global__ void k_rows_to_off(const uint32_t* __restrict__ rows, uint64_t n, uint32_t nV,
                              uint64_t* __restrict__ off) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = rows[i];
    const int64_t rprev = i == 0 ? -1 : (int64_t)rows[i - 1];
    for (int64_t q = rprev + 1; q <= (int64_t)r; ++q) off[q] = i;
    if (i == n - 1)
      for (uint64_t q = (uint64_t)r + 1; q <= nV; ++q) off[q] = n;
  }
}

__global__ void k_fill_u64(uint64_t* p, uint64_t n, uint64_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_transpose_keys(const uint64_t* __restrict__ in, uint64_t n, int shift,
                                 uint64_t* __restrict__ out) {
  const uint64_t mask = (1ULL << shift) - 1;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t k = in[i];
    out[i] = ((k & mask) << shift) | (k >> shift);
  }
}

__global__ void k_outdeg(const uint64_t* __restrict__ off, uint32_t nV, int32_t* __restrict__ d) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV;
       v += (uint64_t)gridDim.x * blockDim.x)
    d[v] = (int32_t)(off[v + 1] - off[v]);
}

__global__ void k_iota_u64(uint64_t* p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = i;
}

__global__ void k_gather_u32(const uint32_t* __restrict__ src, const uint64_t* __restrict__ perm,
                             uint64_t n, uint32_t* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = src[perm[i]];
}

__global__ void k_scatter_pos(const uint64_t* __restrict__ perm, uint64_t n, uint64_t* __restrict__ inv) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    inv[perm[i]] = i;
}

__global__ void k_gather_u64(const uint64_t* __restrict__ src, const uint64_t* __restrict__ perm,
                             uint64_t n, uint64_t* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = src[perm[i]];
}

// src of every CSR entry
__global__ void k_csr_src(const uint64_t* __restrict__ off, uint32_t nV, uint64_t nE,
                          uint32_t* __restrict__ src) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nE;
       e += (uint64_t)gridDim.x * blockDim.x)
    src[e] = row_of(off, nV, e);
}

// CSC position -> CSR position when CSR rows are sorted: binary search of the
// row of the source for the destination.
__global__ void k_ceid_sorted(const uint64_t* __restrict__ coff, const uint32_t* __restrict__ cidx,
                              const uint64_t* __restrict__ off, const uint32_t* __restrict__ idx,
                              uint32_t nV, uint64_t nE, uint64_t* __restrict__ ceid) {
  for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < nE;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = row_of(coff, nV, p);
    const uint32_t u = cidx[p];
    uint64_t lo = off[u], hi = off[u + 1];
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (idx[mid] < v) lo = mid + 1; else hi = mid;
    }
    ceid[p] = lo;
  }
}

__global__ void k_sssp_weights(const uint64_t* __restrict__ off, const uint32_t* __restrict__ idx,
                               uint32_t nV, uint64_t nE, uint64_t seed, uint32_t* __restrict__ w) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < nE;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = row_of(off, nV, e), v = idx[e];
    const uint32_t a = u < v ? u : v, b = u < v ? v : u;
    w[e] = 1u + (uint32_t)(mix64(seed * GRE_K0 + (uint64_t)a * GRE_K1 + (uint64_t)b * GRE_K2 + GRE_K3) % 255ULL);
  }
}

__global__ void k_gather_w(const uint32_t* __restrict__ w, const uint64_t* __restrict__ ceid,
                           uint64_t n, uint32_t* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = w[ceid[i]];
}

// tile_rb[t] = number of rows r < nV with off[r] < t*te
template <class Off>
__global__ void k_tiles(const Off* __restrict__ off, uint32_t nV, uint64_t ntiles, uint32_t te,
                        uint32_t* __restrict__ rb) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t <= ntiles;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t e = t * (uint64_t)te;
    uint32_t lo = 0, hi = nV;  // first r in [0,nV) with off[r] >= e, or nV
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (off[mid] < e) lo = mid + 1; else hi = mid;
    }
    rb[t] = (t == ntiles) ? nV : lo;
  }
}

unsigned grid_cap(uint64_t n) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

template <typename K>
void sort_keys(K* in, K* out, size_t n, unsigned end_bit, hipStream_t s) {
  size_t tb = 0;
  GRE_HIP(rocprim::radix_sort_keys(nullptr, tb, in, out, n, 0, end_bit, s));
  void* tmp = scratch(0, tb);
  GRE_HIP(rocprim::radix_sort_keys(tmp, tb, in, out, n, 0, end_bit, s));
}

template <typename K, typename V>
void sort_pairs(K* kin, K* kout, V* vin, V* vout, size_t n, unsigned end_bit, hipStream_t s) {
  size_t tb = 0;
  GRE_HIP(rocprim::radix_sort_pairs(nullptr, tb, kin, kout, vin, vout, n, 0, end_bit, s));
  void* tmp = scratch(0, tb);
  GRE_HIP(rocprim::radix_sort_pairs(tmp, tb, kin, kout, vin, vout, n, 0, end_bit, s));
}

uint64_t unique_keys(uint64_t* in, uint64_t* out, size_t n, hipStream_t s) {
  size_t tb = 0;
  uint64_t* cnt = static_cast<uint64_t*>(scratch(1, sizeof(uint64_t)));
  GRE_HIP(rocprim::unique(nullptr, tb, in, out, cnt, n, rocprim::equal_to<uint64_t>(), s));
  void* tmp = scratch(0, tb);
  GRE_HIP(rocprim::unique(tmp, tb, in, out, cnt, n, rocprim::equal_to<uint64_t>(), s));
  uint64_t h = 0;
  GRE_HIP(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
  GRE_HIP(hipStreamSynchronize(s));
  return h;
}

unsigned bits_for(uint32_t nV) {
  unsigned b = 1;
  while (b < 32 && (1ULL << b) < nV) ++b;
  return b;
}

void finish_graph(gre_graph* g, hipStream_t s) {
  g->outdeg.alloc(sizeof(int32_t) * (size_t)g->nV);
  hipLaunchKernelGGL(k_outdeg, dim3(grid_cap(g->nV)), dim3(kBlock), 0, s, g->off.as<uint64_t>(), g->nV,
                     g->outdeg.as<int32_t>());
  GRE_HIP(hipGetLastError());
  GRE_HIP(hipStreamSynchronize(s));
}

// off[0..nV] for an empty edge set
void zero_offsets(gre::DevBuf& off, uint32_t nV, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_cap(nV + 1)), dim3(kBlock), 0, s, off.as<uint64_t>(),
                     (uint64_t)nV + 1, 0ULL);
  GRE_HIP(hipGetLastError());
}

}  // namespace

void scan_excl_u64(const uint64_t* in, uint64_t* out, size_t n, hipStream_t s) {
  size_t tb = 0;
  GRE_HIP(rocprim::exclusive_scan(nullptr, tb, in, out, (uint64_t)0, n, rocprim::plus<uint64_t>(), s));
  void* tmp = scratch(2, tb);
  GRE_HIP(rocprim::exclusive_scan(tmp, tb, in, out, (uint64_t)0, n, rocprim::plus<uint64_t>(), s));
}

void scan_excl_u32_to_u64(const uint32_t* in, uint64_t* out, size_t n, hipStream_t s) {
  size_t tb = 0;
  GRE_HIP(rocprim::exclusive_scan(nullptr, tb, in, out, (uint64_t)0, n, rocprim::plus<uint64_t>(), s));
  void* tmp = scratch(2, tb);
  GRE_HIP(rocprim::exclusive_scan(tmp, tb, in, out, (uint64_t)0, n, rocprim::plus<uint64_t>(), s));
}

const uint64_t* graph_ceid(gre_graph* g) {
  if (g->ceid.p) return g->ceid.as<uint64_t>();
  hipStream_t s = stream();
  g->ceid.alloc(sizeof(uint64_t) * (size_t)g->nE);
  if (g->nE == 0) return g->ceid.as<uint64_t>();
  if (g->rows_sorted) {
    hipLaunchKernelGGL(k_ceid_sorted, dim3(grid_cap(g->nE)), dim3(kBlock), 0, s, g->coff.as<uint64_t>(),
                       g->cidx.as<uint32_t>(), g->off.as<uint64_t>(), g->idx.as<uint32_t>(), g->nV, g->nE,
                       g->ceid.as<uint64_t>());
    GRE_HIP(hipGetLastError());
  } else {
    // stable sort of CSR positions by destination gives the CSC order
    DevBuf kin, kout, vin;
    kin.alloc(sizeof(uint32_t) * g->nE);
    kout.alloc(sizeof(uint32_t) * g->nE);
    vin.alloc(sizeof(uint64_t) * g->nE);
    GRE_HIP(hipMemcpyAsync(kin.p, g->idx.p, sizeof(uint32_t) * g->nE, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_iota_u64, dim3(grid_cap(g->nE)), dim3(kBlock), 0, s, vin.as<uint64_t>(), g->nE);
    sort_pairs(kin.as<uint32_t>(), kout.as<uint32_t>(), vin.as<uint64_t>(), g->ceid.as<uint64_t>(),
               g->nE, bits_for(g->nV), s);
  }
  GRE_HIP(hipStreamSynchronize(s));
  return g->ceid.as<uint64_t>();
}

const uint32_t* graph_w_csc(gre_graph* g) {
  if (g->w_u32_csc.p) return g->w_u32_csc.as<uint32_t>();
  GRE_CHECK(g->w_u32.p, GRE_DEV_NULL, "graph has no edge weights");
  const uint64_t* ceid = graph_ceid(g);
  hipStream_t s = stream();
  g->w_u32_csc.alloc(sizeof(uint32_t) * (size_t)g->nE);
  hipLaunchKernelGGL(k_gather_w, dim3(grid_cap(g->nE)), dim3(kBlock), 0, s, g->w_u32.as<uint32_t>(), ceid,
                     g->nE, g->w_u32_csc.as<uint32_t>());
  GRE_HIP(hipGetLastError());
  return g->w_u32_csc.as<uint32_t>();
}

namespace {
__global__ void k_iota_u32(uint32_t* p, uint32_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = (uint32_t)i;
}
__global__ void k_outdeg_u32(const int32_t* __restrict__ d, uint32_t n, uint32_t* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (uint32_t)d[i];
}
__global__ void k_indeg_u32(const uint64_t* __restrict__ coff, uint32_t n, uint32_t* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (uint32_t)(coff[i + 1] - coff[i]);
}
// cid[order[i]] = i for the nsrc vertices with out-edges, ~0 for the others
__global__ void k_rank(const uint32_t* __restrict__ order, const uint32_t* __restrict__ deg_sorted, uint32_t n,
                       uint32_t* __restrict__ cid, uint32_t* __restrict__ nsrc) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const bool has = deg_sorted[i] > 0;
    cid[order[i]] = has ? (uint32_t)i : 0xffffffffu;
    if (has && (i + 1 == n || deg_sorted[i + 1] == 0)) *nsrc = (uint32_t)(i + 1);
  }
}
__global__ void k_map_idx(const uint32_t* __restrict__ idx, const uint32_t* __restrict__ cid, uint64_t n,
                          uint32_t* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = cid[idx[i]];
}
}  // namespace

// Ranks the vertices by descending degree (ties by vertex id: the sort is
// stable); id[v] = rank for vertices of non-zero degree, ~0 for the others;
// idx_c[p] = id[idx[p]]; returns the number of ranked vertices.
static uint32_t compact_ids(gre_graph* g, bool out_degree, const uint32_t* idx, DevBuf& id, DevBuf& idx_c) {
  hipStream_t s = stream();
  const uint32_t n = g->nV;
  DevBuf keys, keys_sorted, vals, order, cnt;
  keys.alloc(4 * (size_t)n);
  keys_sorted.alloc(4 * (size_t)n);
  vals.alloc(4 * (size_t)n);
  order.alloc(4 * (size_t)n);
  cnt.alloc(8);
  GRE_HIP(hipMemsetAsync(cnt.p, 0, 8, s));
  if (out_degree)
    hipLaunchKernelGGL(k_outdeg_u32, dim3(grid_cap(n)), dim3(kBlock), 0, s, g->outdeg.as<int32_t>(), n,
                       keys.as<uint32_t>());
  else
    hipLaunchKernelGGL(k_indeg_u32, dim3(grid_cap(n)), dim3(kBlock), 0, s, g->coff.as<uint64_t>(), n,
                       keys.as<uint32_t>());
  hipLaunchKernelGGL(k_iota_u32, dim3(grid_cap(n)), dim3(kBlock), 0, s, vals.as<uint32_t>(), n);
  GRE_HIP(hipGetLastError());
  size_t tb = 0;
  GRE_HIP(rocprim::radix_sort_pairs_desc(nullptr, tb, keys.as<uint32_t>(), keys_sorted.as<uint32_t>(),
                                         vals.as<uint32_t>(), order.as<uint32_t>(), (size_t)n, 0, 32, s));
  void* tmp = scratch(0, tb);
  GRE_HIP(rocprim::radix_sort_pairs_desc(tmp, tb, keys.as<uint32_t>(), keys_sorted.as<uint32_t>(),
                                         vals.as<uint32_t>(), order.as<uint32_t>(), (size_t)n, 0, 32, s));
  id.alloc(4 * (size_t)n);
  hipLaunchKernelGGL(k_rank, dim3(grid_cap(n)), dim3(kBlock), 0, s, order.as<uint32_t>(), keys_sorted.as<uint32_t>(), n,
                     id.as<uint32_t>(), cnt.as<uint32_t>());
  idx_c.alloc(4 * (size_t)g->nE);
  if (g->nE)
    hipLaunchKernelGGL(k_map_idx, dim3(grid_cap(g->nE)), dim3(kBlock), 0, s, idx, id.as<uint32_t>(), g->nE,
                       idx_c.as<uint32_t>());
  GRE_HIP(hipGetLastError());
  uint32_t ns = 0;
  GRE_HIP(hipMemcpyAsync(&ns, cnt.p, 4, hipMemcpyDeviceToHost, s));
  GRE_HIP(hipStreamSynchronize(s));
  return ns;
}

void graph_compact_sources(gre_graph* g) {
  if (g->cid.p) return;
  g->nsrc = compact_ids(g, true, g->cidx.as<uint32_t>(), g->cid, g->cidx_c);
}

void graph_compact_targets(gre_graph* g) {
  if (g->did.p) return;
  g->ntgt = compact_ids(g, false, g->idx.as<uint32_t>(), g->did, g->idx_c);
}

gre_graph* graph_reversed(gre_graph* g) {
  if (g->rev) return g->rev.get();
  hipStream_t s = stream();
  auto r = std::make_unique<gre_graph>();
  r->nV = g->nV;
  r->nE = g->nE;
  r->off.view(g->coff);
  r->idx.view(g->cidx);
  r->coff.view(g->off);
  r->cidx.view(g->idx);
  r->outdeg.alloc(sizeof(int32_t) * (size_t)g->nV);
  hipLaunchKernelGGL(k_outdeg, dim3(grid_cap(g->nV)), dim3(kBlock), 0, s, g->coff.as<uint64_t>(), g->nV,
                     r->outdeg.as<int32_t>());
  GRE_HIP(hipGetLastError());
  GRE_HIP(hipStreamSynchronize(s));
  g->rev = std::move(r);
  return g->rev.get();
}

const uint32_t* graph_tiles(gre_graph* g, bool csc, uint32_t te, uint64_t* ntiles) {
  uint64_t nt = (g->nE + te - 1) / te;
  if (nt == 0) nt = 1;
  *ntiles = nt;
  auto key = std::make_pair(csc ? 1 : 0, te);
  auto it = g->tiles.find(key);
  if (it != g->tiles.end()) return it->second.as<uint32_t>();
  DevBuf& b = g->tiles[key];
  b.alloc(sizeof(uint32_t) * (nt + 1));
  build_tile_rows((csc ? g->coff : g->off).as<uint64_t>(), g->nV, nt, te, b.as<uint32_t>(), stream());
  return b.as<uint32_t>();
}

void build_tile_rows(const uint64_t* off, uint32_t nrows, uint64_t ntiles, uint32_t te, uint32_t* rb,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_tiles<uint64_t>, dim3(grid_cap(ntiles + 1)), dim3(kBlock), 0, s, off, nrows, ntiles, te, rb);
  GRE_HIP(hipGetLastError());
}

void build_tile_rows(const uint32_t* off, uint32_t nrows, uint64_t ntiles, uint32_t te, uint32_t* rb,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_tiles<uint32_t>, dim3(grid_cap(ntiles + 1)), dim3(kBlock), 0, s, off, nrows, ntiles, te, rb);
  GRE_HIP(hipGetLastError());
}

}  // namespace gre

using namespace gre;

extern "C" {

int gre_graph_rmat(gre_graph** out, int scale, int ef, uint64_t seed, int flags) {
  return guard([&] {
    GRE_CHECK(out, GRE_DEV_NULL, "null out pointer");
    GRE_CHECK(flags >= 0 && flags <= (GRE_RMAT_SYM | GRE_RMAT_SCRAMBLE), GRE_DEV_NULL, "unknown R-MAT flags");
    const int sym = flags & GRE_RMAT_SYM;
    GRE_CHECK(scale >= 1 && scale <= 30, GRE_DEV_NULL, "scale must be in [1,30]");
    GRE_CHECK(ef >= 1 && ef <= 1024, GRE_DEV_NULL, "edge factor must be in [1,1024]");
    hipStream_t s = stream();
    auto* g = new gre_graph();
    try {
      g->nV = 1u << scale;
      const uint64_t ngen = (uint64_t)ef << scale;
      const uint64_t nk = ngen * (sym ? 2 : 1);
      const unsigned kbits = 2 * scale;
      DevBuf a, b;
      a.alloc(sizeof(uint64_t) * nk);
      b.alloc(sizeof(uint64_t) * nk);
      hipLaunchKernelGGL(k_rmat_keys, dim3(grid_cap(ngen)), dim3(kBlock), 0, s, scale, seed, ngen, sym,
                         (flags & GRE_RMAT_SCRAMBLE) ? 1 : 0, a.as<uint64_t>());
      GRE_HIP(hipGetLastError());
      sort_keys(a.as<uint64_t>(), b.as<uint64_t>(), nk, kbits, s);
      uint64_t n = unique_keys(b.as<uint64_t>(), a.as<uint64_t>(), nk, s);
      // drop the self-loop marker (the largest key) if present
      if (n > 0) {
        uint64_t last = 0;
        GRE_HIP(hipMemcpyAsync(&last, a.as<uint64_t>() + (n - 1), 8, hipMemcpyDeviceToHost, s));
        GRE_HIP(hipStreamSynchronize(s));
        if (last == ((1ULL << kbits) - 1)) --n;
      }
      g->nE = n;
      g->rows_sorted = true;
      g->off.alloc(sizeof(uint64_t) * ((size_t)g->nV + 1));
      g->idx.alloc(sizeof(uint32_t) * n);
      g->coff.alloc(sizeof(uint64_t) * ((size_t)g->nV + 1));
      g->cidx.alloc(sizeof(uint32_t) * n);
      if (n == 0) {
        zero_offsets(g->off, g->nV, s);
        zero_offsets(g->coff, g->nV, s);
      } else {
        hipLaunchKernelGGL(k_keys_to_csr, dim3(grid_cap(n)), dim3(kBlock), 0, s, a.as<uint64_t>(), n, scale,
                           g->nV, g->off.as<uint64_t>(), g->idx.as<uint32_t>());
        GRE_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_transpose_keys, dim3(grid_cap(n)), dim3(kBlock), 0, s, a.as<uint64_t>(), n, scale,
                           b.as<uint64_t>());
        GRE_HIP(hipGetLastError());
        sort_keys(b.as<uint64_t>(), a.as<uint64_t>(), n, kbits, s);
        hipLaunchKernelGGL(k_keys_to_csr, dim3(grid_cap(n)), dim3(kBlock), 0, s, a.as<uint64_t>(), n, scale,
                           g->nV, g->coff.as<uint64_t>(), g->cidx.as<uint32_t>());
        GRE_HIP(hipGetLastError());
      }
      GRE_HIP(hipStreamSynchronize(s));
      finish_graph(g, s);
    } catch (...) {
      delete g;
      throw;
    }
    *out = g;
  });
}

int gre_graph_from_edges(gre_graph** out, uint32_t nV, uint64_t nE, const uint32_t* src,
                         const uint32_t* dst) {
  return guard([&] {
    GRE_CHECK(out, GRE_DEV_NULL, "null out pointer");
    GRE_CHECK(nE == 0 || (src && dst), GRE_DEV_NULL, "null edge arrays");
    GRE_CHECK(nV > 0, GRE_DEV_NULL, "nV must be > 0");
    for (uint64_t i = 0; i < nE; ++i)
      GRE_CHECK(src[i] < nV && dst[i] < nV, GRE_DEV_NULL,
                "edge " + std::to_string(i) + " has a vertex id >= nV");
    hipStream_t s = stream();
    auto* g = new gre_graph();
    try {
      g->nV = nV;
      g->nE = nE;
      g->off.alloc(sizeof(uint64_t) * ((size_t)nV + 1));
      g->coff.alloc(sizeof(uint64_t) * ((size_t)nV + 1));
      g->idx.alloc(sizeof(uint32_t) * nE);
      g->cidx.alloc(sizeof(uint32_t) * nE);
      g->ceid.alloc(sizeof(uint64_t) * nE);
      if (nE == 0) {
        zero_offsets(g->off, nV, s);
        zero_offsets(g->coff, nV, s);
      } else {
        DevBuf dsrc, ddst, keys, perm_in, perm1, perm2, inv;
        dsrc.alloc(4 * nE); ddst.alloc(4 * nE); keys.alloc(4 * nE);
        perm_in.alloc(8 * nE); perm1.alloc(8 * nE); perm2.alloc(8 * nE); inv.alloc(8 * nE);
        GRE_HIP(hipMemcpyAsync(dsrc.p, src, 4 * nE, hipMemcpyHostToDevice, s));
        GRE_HIP(hipMemcpyAsync(ddst.p, dst, 4 * nE, hipMemcpyHostToDevice, s));
        const unsigned vb = bits_for(nV);
        hipLaunchKernelGGL(k_iota_u64, dim3(grid_cap(nE)), dim3(kBlock), 0, s, perm_in.as<uint64_t>(), nE);
        // CSR: stable sort of input positions by source
        sort_pairs(dsrc.as<uint32_t>(), keys.as<uint32_t>(), perm_in.as<uint64_t>(), perm1.as<uint64_t>(),
                   nE, vb, s);
        hipLaunchKernelGGL(k_rows_to_off, dim3(grid_cap(nE)), dim3(kBlock), 0, s, keys.as<uint32_t>(), nE, nV,
                           g->off.as<uint64_t>());
        hipLaunchKernelGGL(k_gather_u32, dim3(grid_cap(nE)), dim3(kBlock), 0, s, ddst.as<uint32_t>(),
                           perm1.as<uint64_t>(), nE, g->idx.as<uint32_t>());
        hipLaunchKernelGGL(k_scatter_pos, dim3(grid_cap(nE)), dim3(kBlock), 0, s, perm1.as<uint64_t>(), nE,
                           inv.as<uint64_t>());
        // CSC: stable sort of input positions by destination
        sort_pairs(ddst.as<uint32_t>(), keys.as<uint32_t>(), perm_in.as<uint64_t>(), perm2.as<uint64_t>(),
                   nE, vb, s);
        hipLaunchKernelGGL(k_rows_to_off, dim3(grid_cap(nE)), dim3(kBlock), 0, s, keys.as<uint32_t>(), nE, nV,
                           g->coff.as<uint64_t>());
        hipLaunchKernelGGL(k_gather_u32, dim3(grid_cap(nE)), dim3(kBlock), 0, s, dsrc.as<uint32_t>(),
                           perm2.as<uint64_t>(), nE, g->cidx.as<uint32_t>());
        hipLaunchKernelGGL(k_gather_u64, dim3(grid_cap(nE)), dim3(kBlock), 0, s, inv.as<uint64_t>(),
                           perm2.as<uint64_t>(), nE, g->ceid.as<uint64_t>());
        GRE_HIP(hipGetLastError());
        GRE_HIP(hipStreamSynchronize(s));
      }
      finish_graph(g, s);
    } catch (...) {
      delete g;
      throw;
    }
    *out = g;
  });
}

int gre_graph_from_csr(gre_graph** out, uint32_t nV, uint64_t nE, const uint64_t* off,
                       const uint32_t* idx) {
  return guard([&] {
    GRE_CHECK(out && off && (nE == 0 || idx), GRE_DEV_NULL, "null argument");
    GRE_CHECK(nV > 0, GRE_DEV_NULL, "nV must be > 0");
    GRE_CHECK(off[0] == 0 && off[nV] == nE, GRE_DEV_NULL, "offsets must start at 0 and end at nE");
    bool sorted = true;
    for (uint32_t v = 0; v < nV; ++v) {
      GRE_CHECK(off[v] <= off[v + 1], GRE_DEV_NULL, "offsets must be non-decreasing");
      for (uint64_t e = off[v]; e < off[v + 1]; ++e) {
        GRE_CHECK(idx[e] < nV, GRE_DEV_NULL, "vertex id >= nV in idx");
        if (e > off[v] && idx[e] <= idx[e - 1]) sorted = false;
      }
    }
    hipStream_t s = stream();
    auto* g = new gre_graph();
    try {
      g->nV = nV;
      g->nE = nE;
      g->rows_sorted = sorted;
      g->off.alloc(sizeof(uint64_t) * ((size_t)nV + 1));
      g->coff.alloc(sizeof(uint64_t) * ((size_t)nV + 1));
      g->idx.alloc(sizeof(uint32_t) * nE);
      g->cidx.alloc(sizeof(uint32_t) * nE);
      GRE_HIP(hipMemcpyAsync(g->off.p, off, 8 * ((size_t)nV + 1), hipMemcpyHostToDevice, s));
      if (nE == 0) {
        zero_offsets(g->coff, nV, s);
      } else {
        GRE_HIP(hipMemcpyAsync(g->idx.p, idx, 4 * nE, hipMemcpyHostToDevice, s));
        DevBuf srcs, keys, perm;
        srcs.alloc(4 * nE); keys.alloc(4 * nE); perm.alloc(8 * nE);
        hipLaunchKernelGGL(k_csr_src, dim3(grid_cap(nE)), dim3(kBlock), 0, s, g->off.as<uint64_t>(), nV, nE,
                           srcs.as<uint32_t>());
        // stable sort of CSR positions by destination
        DevBuf pin;
        pin.alloc(8 * nE);
        hipLaunchKernelGGL(k_iota_u64, dim3(grid_cap(nE)), dim3(kBlock), 0, s, pin.as<uint64_t>(), nE);
        sort_pairs(g->idx.as<uint32_t>(), keys.as<uint32_t>(), pin.as<uint64_t>(), perm.as<uint64_t>(), nE,
                   bits_for(nV), s);
        hipLaunchKernelGGL(k_rows_to_off, dim3(grid_cap(nE)), dim3(kBlock), 0, s, keys.as<uint32_t>(), nE, nV,
                           g->coff.as<uint64_t>());
        hipLaunchKernelGGL(k_gather_u32, dim3(grid_cap(nE)), dim3(kBlock), 0, s, srcs.as<uint32_t>(),
                           perm.as<uint64_t>(), nE, g->cidx.as<uint32_t>());
        GRE_HIP(hipGetLastError());
        g->ceid = std::move(perm);
        GRE_HIP(hipStreamSynchronize(s));
      }
      finish_graph(g, s);
    } catch (...) {
      delete g;
      throw;
    }
    *out = g;
  });
}

int gre_graph_destroy(gre_graph* g) {
  return guard([&] {
    if (!g) return;
    GRE_HIP(hipStreamSynchronize(stream()));
    delete g;
  });
}

int gre_graph_info(const gre_graph* g, uint32_t* nV, uint64_t* nE) {
  return guard([&] {
    GRE_CHECK(g, GRE_DEV_NULL, "null graph");
    if (nV) *nV = g->nV;
    if (nE) *nE = g->nE;
  });
}

int gre_graph_download(const gre_graph* gc, uint64_t* off, uint32_t* idx, uint64_t* coff, uint32_t* cidx,
                       uint64_t* ceid, int32_t* outdeg) {
  return guard([&] {
    GRE_CHECK(gc, GRE_DEV_NULL, "null graph");
    gre_graph* g = const_cast<gre_graph*>(gc);
    hipStream_t s = stream();
    const size_t nv1 = (size_t)g->nV + 1;
    if (off) GRE_HIP(hipMemcpyAsync(off, g->off.p, 8 * nv1, hipMemcpyDeviceToHost, s));
    if (coff) GRE_HIP(hipMemcpyAsync(coff, g->coff.p, 8 * nv1, hipMemcpyDeviceToHost, s));
    if (idx && g->nE) GRE_HIP(hipMemcpyAsync(idx, g->idx.p, 4 * g->nE, hipMemcpyDeviceToHost, s));
    if (cidx && g->nE) GRE_HIP(hipMemcpyAsync(cidx, g->cidx.p, 4 * g->nE, hipMemcpyDeviceToHost, s));
    if (ceid && g->nE) {
      const uint64_t* d = graph_ceid(g);
      GRE_HIP(hipMemcpyAsync(ceid, d, 8 * g->nE, hipMemcpyDeviceToHost, s));
    }
    if (outdeg) GRE_HIP(hipMemcpyAsync(outdeg, g->outdeg.p, 4 * (size_t)g->nV, hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
  });
}

int gre_graph_set_weights_u32(gre_graph* g, const uint32_t* w, uint64_t seed) {
  return guard([&] {
    GRE_CHECK(g, GRE_DEV_NULL, "null graph");
    hipStream_t s = stream();
    g->w_u32.alloc(4 * (size_t)g->nE);
    g->w_u32_csc.release();
    if (g->nE == 0) return;
    if (w) {
      GRE_HIP(hipMemcpyAsync(g->w_u32.p, w, 4 * g->nE, hipMemcpyHostToDevice, s));
    } else {
      hipLaunchKernelGGL(k_sssp_weights, dim3(grid_cap(g->nE)), dim3(kBlock), 0, s, g->off.as<uint64_t>(),
                         g->idx.as<uint32_t>(), g->nV, g->nE, seed, g->w_u32.as<uint32_t>());
      GRE_HIP(hipGetLastError());
    }
    GRE_HIP(hipStreamSynchronize(s));
  });
}

int gre_graph_get_weights_u32(const gre_graph* g, uint32_t* w) {
  return guard([&] {
    GRE_CHECK(g && w, GRE_DEV_NULL, "null argument");
    GRE_CHECK(g->w_u32.p, GRE_DEV_NULL, "graph has no edge weights");
    if (g->nE == 0) return;
    GRE_HIP(hipMemcpyAsync(w, g->w_u32.p, 4 * g->nE, hipMemcpyDeviceToHost, stream()));
    GRE_HIP(hipStreamSynchronize(stream()));
  });
}

}  // extern "C"
