// Source-range partition of the CSC, for pulls whose gathers stay inside one
// XCD's L2.
//
// MI355X has 8 XCDs, each with its own 4 MB L2, and workgroups are dispatched
// to them round-robin by workgroup id.  A plain pull gathers from the whole
// source array on every XCD, so each L2 holds the same hot sources and misses
// on the rest.  Here the hub-sorted compact source ids are cut into ranges:
// the H hottest sources form one shared range, small enough to sit in every
// L2, whose entries are cut by rows into 8 chunks (chunk c on XCD c); the
// other sources form R tail ranges of equal edge count, range r owned by XCD
// class r % 8.  The CSC entries are regrouped by (class, range, row).  Each run
// of one row's entries from one range is a segment; the pull reduces segments,
// and a second tile-segmented pull over per-row lists of segment ids sums each
// row's segments in a fixed order.
#include "graph.hpp"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>

namespace gre {
namespace {

constexpr unsigned kB = 256;

unsigned gcap(uint64_t n) {
  uint64_t g = (n + kB - 1) / kB;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

// degc[cid[v]] = out-degree of v, for the vertices with out-edges
__global__ void k_deg_compact(const int32_t* __restrict__ outdeg, const uint32_t* __restrict__ cid, uint32_t nV,
                              uint64_t* __restrict__ degc) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV; v += (uint64_t)gridDim.x * blockDim.x) {
    const int32_t d = outdeg[v];
    if (d > 0) degc[cid[v]] = (uint64_t)d;
  }
}
// the same for compact target ids: in-degrees from the CSC offsets
__global__ void k_indeg_compact(const uint64_t* __restrict__ coff, const uint32_t* __restrict__ did, uint32_t nV,
                                uint64_t* __restrict__ degc) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t d = coff[v + 1] - coff[v];
    if (d > 0) degc[did[v]] = d;
  }
}

// Tail range bounds: bnd[r] = first compact id c in [H, nsrc] with
// ex[c] >= ex[H] + r * (E - ex[H]) / R (ex: exclusive prefix of the compact
// out-degrees, ex[nsrc] = E); one thread per r in [0, R]
__global__ void k_bounds(const uint64_t* __restrict__ ex, uint32_t H, uint32_t nsrc, uint64_t E, int R,
                         uint32_t* __restrict__ bnd) {
  const int r = (int)threadIdx.x;
  if (r > R) return;
  const uint64_t base = ex[H];
  const uint64_t target = base + (uint64_t)r * (E - base) / (uint64_t)R;
  uint32_t lo = H, hi = nsrc;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (ex[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  bnd[r] = lo;
}

// cnt[v] += entries of row v whose compact source is < H: one thread per 8
// consecutive CSC entries (rows found inside the entries' 2048-entry tile, rb
// from graph_tiles), one atomic per run of a row
__global__ void k_hub_rowcnt(const uint64_t* __restrict__ coff, const uint32_t* __restrict__ rb, uint32_t te,
                             const uint32_t* __restrict__ cidx_c, uint64_t E, uint32_t H,
                             unsigned long long* __restrict__ cnt) {
  for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c * 8 < E; c += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t e0 = c * 8, e1 = e0 + 8 < E ? e0 + 8 : E;
    const uint64_t t = e0 / te;
    uint32_t lo = rb[t] > 0 ? rb[t] - 1 : 0, hi = rb[t + 1];
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (coff[mid] <= e0) lo = mid;
      else hi = mid;
    }
    uint32_t r = lo;
    uint64_t rend = coff[r + 1];
    unsigned long long run = 0;
    for (uint64_t e = e0; e < e1; ++e) {
      if (rend <= e) {
        if (run) atomicAdd(cnt + r, run);
        run = 0;
        do {
          ++r;
          rend = coff[r + 1];
        } while (rend <= e);
      }
      run += cidx_c[e] < H ? 1ull : 0ull;
    }
    if (run) atomicAdd(cnt + r, run);
  }
}

// rb[c] = first row v in [0, nV] with ex[v] >= c * ex[nV] / 8, c in [0, 8]
__global__ void k_row_bounds(const uint64_t* __restrict__ ex, uint32_t nV, uint32_t* __restrict__ rb) {
  const int c = (int)threadIdx.x;
  if (c > 8) return;
  const uint64_t target = (uint64_t)c * ex[nV] / 8;
  uint32_t lo = 0, hi = nV;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (ex[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  rb[c] = lo;
}

// Per CSC entry e: the sort key and the value (row << 32 | compact source).
// Sources < H form the shared range, cut by rows into 8 chunks of equal entry
// count (chunk c runs on XCD c); the other sources fall into R tail ranges,
// range r on XCD r % 8.  key = class * (1 + R/8) + (0 for the shared chunk,
// 1 + r/8 for tail range r).  The row is found inside the entry's 2048-entry
// tile (rb: graph_tiles of the CSC).
__global__ void k_xkeys(const uint64_t* __restrict__ coff, const uint32_t* __restrict__ rb, uint32_t te,
                        const uint32_t* __restrict__ cidx_c, uint64_t E, const uint32_t* __restrict__ bnd, int R,
                        uint32_t H, const uint32_t* __restrict__ hub_rows, uint16_t* __restrict__ key,
                        uint64_t* __restrict__ val) {
  __shared__ uint32_t sb[257];
  __shared__ uint32_t hr[9];
  for (int i = (int)threadIdx.x; i <= R; i += (int)blockDim.x) sb[i] = bnd[i];
  if (threadIdx.x < 9) hr[threadIdx.x] = hub_rows[threadIdx.x];
  __syncthreads();
  const int per = 1 + R / 8;
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E; e += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t t = e / te;
    // largest row r in [lo, hi) with coff[r] <= e
    uint32_t lo = rb[t] > 0 ? rb[t] - 1 : 0, hi = rb[t + 1];
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (coff[mid] <= e) lo = mid;
      else hi = mid;
    }
    const uint32_t s = cidx_c[e];
    int k;
    if (s < H) {
      int c = 0;
      for (int i = 1; i < 8; ++i) c += (hr[i] <= lo) ? 1 : 0;
      k = c * per;
    } else {
      // tail range = (number of inner bounds <= s)
      int a = 1, b = R;
      while (a < b) {
        const int mid = (a + b) >> 1;
        if (sb[mid] <= s) a = mid + 1;
        else b = mid;
      }
      const int r = a - 1;
      k = (r & 7) * per + 1 + (r >> 3);
    }
    key[e] = (uint16_t)k;
    val[e] = ((uint64_t)lo << 32) | s;
  }
}

__global__ void k_xflags(const uint16_t* __restrict__ key, const uint64_t* __restrict__ val, uint64_t E,
                         uint32_t* __restrict__ flag, uint32_t* __restrict__ idx) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < E; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t v = val[i];
    const bool f = i == 0 || key[i] != key[i - 1] || (v >> 32) != (val[i - 1] >> 32);
    flag[i] = f ? 1u : 0u;
    idx[i] = (uint32_t)v;
  }
}

__global__ void k_xsegs(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ segid,
                        const uint64_t* __restrict__ val, uint64_t E, uint32_t* __restrict__ seg_off,
                        uint32_t* __restrict__ seg_row) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < E; i += (uint64_t)gridDim.x * blockDim.x) {
    if (flag[i]) {
      const uint32_t s = segid[i] - 1;
      seg_off[s] = (uint32_t)i;
      seg_row[s] = (uint32_t)(val[i] >> 32);
    }
  }
}

// cs[c] = first entry whose key belongs to class >= c (keys sorted ascending)
__global__ void k_xclass(const uint16_t* __restrict__ key, uint64_t E, int per, uint64_t* __restrict__ cs) {
  const int c = threadIdx.x;
  if (c > 8) return;
  const unsigned target = (unsigned)(c * per);
  uint64_t lo = 0, hi = E;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (key[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  cs[c] = lo;
}

__global__ void k_xpcount(const uint32_t* __restrict__ seg_row, uint64_t nseg, uint32_t* __restrict__ cnt) {
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < nseg; s += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[seg_row[s]], 1u);
}

__global__ void k_gather_u64(const uint64_t* __restrict__ in, const uint32_t* __restrict__ at, uint64_t n,
                             uint64_t* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = in[at[i]];
}

__global__ void k_iota32(uint32_t* __restrict__ o, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    o[i] = (uint32_t)i;
}

// kseg[k] = first segment whose key is >= k, k in [0, K] (keys ascending; a
// key change always starts a segment)
__global__ void k_xkeyseg(const uint16_t* __restrict__ key, const uint32_t* __restrict__ segid, uint64_t E, int K,
                          uint32_t nseg, uint32_t* __restrict__ kseg) {
  for (int k = (int)threadIdx.x; k <= K; k += (int)blockDim.x) {
    uint64_t lo = 0, hi = E;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if ((int)key[mid] < k) lo = mid + 1;
      else hi = mid;
    }
    kseg[k] = lo < E ? segid[lo] - 1 : nseg;
  }
}

// key of segment s: the largest k with kseg[k] <= s (kseg[0] = 0 <= s < kseg[K])
__device__ inline int seg_key(const uint32_t* kseg, int K, uint32_t s) {
  int lo = 0, hi = K;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (kseg[mid] <= s) lo = mid;
    else hi = mid;
  }
  return lo;
}

// pos[s] = position of segment s in the pair lists
__global__ void k_xpos(const uint32_t* __restrict__ pair_list, uint64_t nseg, uint32_t* __restrict__ pos) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < nseg; j += (uint64_t)gridDim.x * blockDim.x)
    pos[pair_list[j]] = (uint32_t)j;
}

// runA[t*K + k] = first segment s of key k with pos[s] >= t*te, or the key's
// end; t in [0, ntiles].  Inside a key, pos grows with s (rows ascending).
__global__ void k_xruna(const uint32_t* __restrict__ kseg, int K, const uint32_t* __restrict__ pos, uint64_t ntiles,
                        uint32_t te, uint32_t* __restrict__ runA) {
  const uint64_t cells = (ntiles + 1) * (uint64_t)K;
  for (uint64_t c = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; c < cells; c += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t t = c / (uint64_t)K;
    const int k = (int)(c % (uint64_t)K);
    const uint64_t target = t * te;
    uint32_t lo = kseg[k], hi = kseg[k + 1];
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if ((uint64_t)pos[mid] < target) lo = mid + 1;
      else hi = mid;
    }
    runA[c] = lo;
  }
}

// runO[t*(K+1) + k] = entries of tile t in keys below k (the LDS offset of key
// k's run); runO[t*(K+1) + K] = the tile's entry count
__global__ void k_xruno(const uint32_t* __restrict__ runA, int K, uint64_t ntiles, uint16_t* __restrict__ runO) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < ntiles; t += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t acc = 0;
    for (int k = 0; k < K; ++k) {
      runO[t * (K + 1) + k] = (uint16_t)acc;
      acc += runA[(t + 1) * K + k] - runA[t * K + k];
    }
    runO[t * (K + 1) + K] = (uint16_t)acc;
  }
}

// slot[j] = LDS slot of pair entry j's segment inside its tile's staged runs
__global__ void k_xslot(const uint32_t* __restrict__ pair_list, uint64_t nseg, uint32_t te,
                        const uint32_t* __restrict__ kseg, int K, const uint32_t* __restrict__ runA,
                        const uint16_t* __restrict__ runO, uint16_t* __restrict__ slot) {
  __shared__ uint32_t ks[kMaxKeys + 1];
  for (int k = (int)threadIdx.x; k <= K; k += (int)blockDim.x) ks[k] = kseg[k];
  __syncthreads();
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < nseg; j += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t s = pair_list[j];
    const uint64_t t = j / te;
    const int k = seg_key(ks, K, s);
    slot[j] = (uint16_t)(runO[t * (K + 1) + k] + (s - runA[t * K + k]));
  }
}

// Per tile: the run-start bitmap with its running counts, and each non-empty
// run's segment base (first segment minus stage offset), runs in key order
__global__ void k_xstage(const uint32_t* __restrict__ runA, const uint16_t* __restrict__ runO, int K,
                         uint64_t ntiles, uint32_t W, uint2* __restrict__ words, uint32_t* __restrict__ base) {
  for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < ntiles; t += (uint64_t)gridDim.x * blockDim.x) {
    uint2* wt = words + t * W;
    uint32_t m = 0;
    for (int k = 0; k < K; ++k) {
      const uint32_t o = runO[t * (K + 1) + k], o1 = runO[t * (K + 1) + k + 1];
      if (o1 > o) {
        wt[o >> 5].x |= 1u << (o & 31);
        base[t * K + m] = runA[t * K + k] - o;
        ++m;
      }
    }
    uint32_t acc = 0;
    for (uint32_t w = 0; w < W; ++w) {
      wt[w].y = acc;
      acc += __popc(wt[w].x);
    }
  }
}

// The row relabelling: rows with more than thr in-edges (the heavy rows:
// R-MAT puts them at the low ids) are spread evenly among the others, which
// keep their order.  Heavy row j goes before light row ins(j) = floor((2j+1)
// L / 2H) (L light and H heavy rows), so light row l lands at l + #{j :
// ins(j) <= l} and heavy row j at ins(j) + j: a merge, so a bijection.  The
// light rows keep their order, so the row pass's writes of x / out-degree by
// compact source id keep the locality they had; the heavy rows no longer fill
// the first tiles.  At R-MAT scale 26 (part_nt, 256 ranges, 4,096 shared):
// 5.13 ms per iteration unrelabelled, 4.73 with heavy = more than 8 times the
// mean in-degree (4.77 at 2 times, 4.78 at 32, 4.74 at 128); a multiplicative
// bijection of all rows gave 5.03 (its scattered compact-id writes cost the
// row pass 2.6 times the write bytes).
constexpr uint64_t kHeavy = 8;
__global__ void k_heavy(uint32_t nV, const uint64_t* __restrict__ coff, uint64_t thr, uint64_t* __restrict__ h) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV; v += (uint64_t)gridDim.x * blockDim.x)
    h[v] = coff[v + 1] - coff[v] > thr ? 1 : 0;
}
__global__ void k_rowperm(uint32_t nV, const uint64_t* __restrict__ h, const uint64_t* __restrict__ hrank,
                          uint64_t H, uint32_t* __restrict__ perm, uint32_t* __restrict__ inv) {
  const uint64_t L = nV - H;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV; v += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t r = v;
    if (H > 0 && L > 0) {
      const uint64_t j = hrank[v];
      if (h[v]) {
        r = (2 * j + 1) * L / (2 * H) + j;
      } else {
        const uint64_t l = v - j;
        const uint64_t n = 2 * H * (l + 1);
        uint64_t before = 0;
        if (n > L) {
          before = (n - L + 2 * L - 1) / (2 * L);
          if (before > H) before = H;
        }
        r = l + before;
      }
    }
    perm[v] = (uint32_t)r;
    inv[r] = (uint32_t)v;
  }
}

// cnt[r] = in-degree of row r's vertex; with the forward's row data in row order
__global__ void k_rowdata(uint32_t nV, const uint32_t* __restrict__ inv, const uint64_t* __restrict__ coff,
                          const int32_t* __restrict__ outdeg, const uint32_t* __restrict__ cid,
                          uint64_t* __restrict__ cnt, int32_t* __restrict__ outdeg_r, uint32_t* __restrict__ cid_r) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < nV; r += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = inv[r];
    cnt[r] = coff[v + 1] - coff[v];
    outdeg_r[r] = outdeg[v];
    cid_r[r] = cid[v];
  }
}

// The CSC with its rows relabelled: entry e of row v moves to coff2[perm[v]] +
// (e - coff[v]), keeping its compact source and recording its old position
// (row found inside the entry's tile of graph_tiles, as in k_xkeys)
__global__ void k_rowmove(const uint64_t* __restrict__ coff, const uint32_t* __restrict__ rb, uint32_t te,
                          const uint32_t* __restrict__ cidx_c, uint64_t E, const uint32_t* __restrict__ perm,
                          const uint64_t* __restrict__ coff2, uint32_t* __restrict__ cidx2,
                          uint32_t* __restrict__ orig) {
  for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < E; e += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t t = e / te;
    uint32_t lo = rb[t] > 0 ? rb[t] - 1 : 0, hi = rb[t + 1];
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (coff[mid] <= e) lo = mid;
      else hi = mid;
    }
    const uint64_t p = coff2[perm[lo]] + (e - coff[lo]);
    cidx2[p] = cidx_c[e];
    orig[p] = (uint32_t)e;
  }
}

__global__ void k_gather_u32(const uint32_t* __restrict__ in, const uint32_t* __restrict__ at, uint64_t n,
                             uint32_t* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = in[at[i]];
}

unsigned bits_of(uint64_t n) {
  unsigned b = 1;
  while (b < 32 && (1ULL << b) < n) ++b;
  return b;
}

}  // namespace

void prune_xparts(gre_graph* g) {
  for (auto it = g->xpart.begin(); it != g->xpart.end();) {
    if (it->second.use_count() == 1) it = g->xpart.erase(it);
    else ++it;
  }
  if (g->rev) prune_xparts(g->rev.get());
}

std::shared_ptr<XPart> graph_xpart(gre_graph* g, int R, uint32_t H, bool csc, bool rows_permuted) {
  GRE_CHECK(R == 8 || R == 16 || R == 32 || R == 64 || R == 128 || R == 256, GRE_DEV_NULL,
            "source ranges must be 8, 16, 32, 64, 128 or 256");
  GRE_CHECK(csc || !rows_permuted, GRE_DEV_NULL, "relabelled rows apply to the CSC partition only");
  // csc: rows are destinations, gathering from compact sources (the forward);
  // otherwise rows are sources, gathering from compact targets (the adjoint)
  if (csc) graph_compact_sources(g);
  else graph_compact_targets(g);
  const uint32_t nsrc = csc ? g->nsrc : g->ntgt;
  const uint64_t* roff = (csc ? g->coff : g->off).as<uint64_t>();
  const uint32_t* ridx = (csc ? g->cidx_c : g->idx_c).as<uint32_t>();
  H = std::min(H, nsrc);
  const auto mkey = std::make_pair(csc ? (rows_permuted ? (R | 1024) : R) : -R, H);
  auto it = g->xpart.find(mkey);
  if (it != g->xpart.end()) return it->second;
  const uint64_t E = g->nE;
  const uint32_t nV = g->nV;
  GRE_CHECK(E > 0 && E < 0xffffffffULL, GRE_DEV_NULL, "the partitioned schedule needs 0 < edges < 2^32");
  hipStream_t s = stream();
  auto xp = std::make_shared<XPart>();
  xp->R = R;
  xp->H = H;
  // the row tiles of the adjacency the build reads: the graph's own, or those
  // of the relabelled CSC
  uint64_t nt2048 = 0;
  const uint32_t* rb2048 = graph_tiles(g, csc, 2048, &nt2048);
  DevBuf coff2, cidx2, orig, rb2;
  if (rows_permuted) {
    xp->permuted = true;
    xp->rowperm.alloc(4 * (size_t)nV);
    xp->rowinv.alloc(4 * (size_t)nV);
    xp->outdeg_r.alloc(4 * (size_t)nV);
    xp->cid_r.alloc(4 * (size_t)nV);
    coff2.alloc(8 * ((size_t)nV + 1));
    cidx2.alloc(4 * E);
    orig.alloc(4 * E);
    {
      // heavy: more than kHeavy times the mean in-degree
      const uint64_t thr = kHeavy * std::max<uint64_t>(E / std::max<uint32_t>(nV, 1), 1);
      DevBuf h, hr;
      h.alloc(8 * ((size_t)nV + 1));
      hr.alloc(8 * ((size_t)nV + 1));
      hipLaunchKernelGGL(k_heavy, dim3(gcap(nV)), dim3(kB), 0, s, nV, g->coff.as<uint64_t>(), thr, h.as<uint64_t>());
      GRE_HIP(hipMemsetAsync(h.as<uint64_t>() + nV, 0, 8, s));
      scan_excl_u64(h.as<uint64_t>(), hr.as<uint64_t>(), (size_t)nV + 1, s);
      uint64_t H = 0;
      GRE_HIP(hipMemcpyAsync(&H, hr.as<uint64_t>() + nV, 8, hipMemcpyDeviceToHost, s));
      GRE_HIP(hipStreamSynchronize(s));
      hipLaunchKernelGGL(k_rowperm, dim3(gcap(nV)), dim3(kB), 0, s, nV, h.as<uint64_t>(), hr.as<uint64_t>(), H,
                         xp->rowperm.as<uint32_t>(), xp->rowinv.as<uint32_t>());
      GRE_HIP(hipGetLastError());
      GRE_HIP(hipStreamSynchronize(s));
    }
    {
      DevBuf cnt;
      cnt.alloc(8 * ((size_t)nV + 1));
      GRE_HIP(hipMemsetAsync(cnt.p, 0, 8 * ((size_t)nV + 1), s));
      hipLaunchKernelGGL(k_rowdata, dim3(gcap(nV)), dim3(kB), 0, s, nV, xp->rowinv.as<uint32_t>(),
                         g->coff.as<uint64_t>(), g->outdeg.as<int32_t>(), g->cid.as<uint32_t>(), cnt.as<uint64_t>(),
                         xp->outdeg_r.as<int32_t>(), xp->cid_r.as<uint32_t>());
      GRE_HIP(hipGetLastError());
      scan_excl_u64(cnt.as<uint64_t>(), coff2.as<uint64_t>(), (size_t)nV + 1, s);
    }
    hipLaunchKernelGGL(k_rowmove, dim3(gcap(E)), dim3(kB), 0, s, g->coff.as<uint64_t>(), rb2048, 2048u,
                       g->cidx_c.as<uint32_t>(), E, xp->rowperm.as<uint32_t>(), coff2.as<uint64_t>(),
                       cidx2.as<uint32_t>(), orig.as<uint32_t>());
    GRE_HIP(hipGetLastError());
    nt2048 = (E + 2047) / 2048;
    rb2.alloc(4 * (nt2048 + 1));
    build_tile_rows(coff2.as<uint64_t>(), nV, nt2048, 2048u, rb2.as<uint32_t>(), s);
    GRE_HIP(hipStreamSynchronize(s));
    roff = coff2.as<uint64_t>();
    ridx = cidx2.as<uint32_t>();
    rb2048 = rb2.as<uint32_t>();
  }

  // tail range bounds from the compact out-degree prefix
  DevBuf bnd, hub_rows;
  bnd.alloc(4 * (size_t)(R + 1));
  hub_rows.alloc(4 * 9);
  {
    DevBuf degc, ex;
    degc.alloc(8 * ((size_t)nsrc + 1));
    ex.alloc(8 * ((size_t)nsrc + 1));
    GRE_HIP(hipMemsetAsync(degc.p, 0, 8 * ((size_t)nsrc + 1), s));
    if (csc)
      hipLaunchKernelGGL(k_deg_compact, dim3(gcap(nV)), dim3(kB), 0, s, g->outdeg.as<int32_t>(),
                         g->cid.as<uint32_t>(), nV, degc.as<uint64_t>());
    else
      hipLaunchKernelGGL(k_indeg_compact, dim3(gcap(nV)), dim3(kB), 0, s, g->coff.as<uint64_t>(),
                         g->did.as<uint32_t>(), nV, degc.as<uint64_t>());
    GRE_HIP(hipGetLastError());
    scan_excl_u64(degc.as<uint64_t>(), ex.as<uint64_t>(), (size_t)nsrc + 1, s);
    hipLaunchKernelGGL(k_bounds, dim3(1), dim3(320), 0, s, ex.as<uint64_t>(), H, nsrc, E, R, bnd.as<uint32_t>());
    GRE_HIP(hipGetLastError());
  }
  // row chunks of the shared range, equal shared-entry counts
  {
    DevBuf cnt, ex;
    cnt.alloc(8 * ((size_t)nV + 1));
    ex.alloc(8 * ((size_t)nV + 1));
    GRE_HIP(hipMemsetAsync(cnt.p, 0, 8 * ((size_t)nV + 1), s));
    hipLaunchKernelGGL(k_hub_rowcnt, dim3(gcap((E + 7) / 8)), dim3(kB), 0, s, roff, rb2048, 2048u, ridx, E, H,
                       cnt.as<unsigned long long>());
    GRE_HIP(hipGetLastError());
    scan_excl_u64(cnt.as<uint64_t>(), ex.as<uint64_t>(), (size_t)nV + 1, s);
    hipLaunchKernelGGL(k_row_bounds, dim3(1), dim3(64), 0, s, ex.as<uint64_t>(), nV, hub_rows.as<uint32_t>());
    GRE_HIP(hipGetLastError());
    GRE_HIP(hipStreamSynchronize(s));
  }

  // keys and values, stable sort by key (rows stay ascending inside a key)
  const int per = 1 + R / 8;
  DevBuf key_s, val_s;
  key_s.alloc(2 * E);
  val_s.alloc(8 * E);
  {
    DevBuf key, val;
    key.alloc(2 * E);
    val.alloc(8 * E);
    const uint32_t te = 2048;
    hipLaunchKernelGGL(k_xkeys, dim3(gcap(E)), dim3(kB), 0, s, roff, rb2048, te, ridx, E, bnd.as<uint32_t>(), R, H,
                       hub_rows.as<uint32_t>(),
                       key.as<uint16_t>(), val.as<uint64_t>());
    GRE_HIP(hipGetLastError());
    // a stable sort of the entries' positions by key, then the values in that
    // order (the positions stay: they carry edge weights to the new order)
    DevBuf pos;
    pos.alloc(4 * E);
    xp->perm.alloc(4 * E);
    hipLaunchKernelGGL(k_iota32, dim3(gcap(E)), dim3(kB), 0, s, pos.as<uint32_t>(), E);
    GRE_HIP(hipGetLastError());
    size_t tb = 0;
    const unsigned end_bit = bits_of((uint64_t)(8 * per));
    GRE_HIP(rocprim::radix_sort_pairs(nullptr, tb, key.as<uint16_t>(), key_s.as<uint16_t>(), pos.as<uint32_t>(),
                                      xp->perm.as<uint32_t>(), (size_t)E, 0, end_bit, s));
    void* tmp = scratch(0, tb);
    GRE_HIP(rocprim::radix_sort_pairs(tmp, tb, key.as<uint16_t>(), key_s.as<uint16_t>(), pos.as<uint32_t>(),
                                      xp->perm.as<uint32_t>(), (size_t)E, 0, end_bit, s));
    hipLaunchKernelGGL(k_gather_u64, dim3(gcap(E)), dim3(kB), 0, s, val.as<uint64_t>(), xp->perm.as<uint32_t>(), E,
                       val_s.as<uint64_t>());
    GRE_HIP(hipGetLastError());
    if (rows_permuted) {
      // perm names positions of the relabelled CSC: take them back to the
      // graph's CSC, where the edge weights are
      hipLaunchKernelGGL(k_gather_u32, dim3(gcap(E)), dim3(kB), 0, s, orig.as<uint32_t>(), xp->perm.as<uint32_t>(),
                         E, pos.as<uint32_t>());
      GRE_HIP(hipGetLastError());
      std::swap(xp->perm, pos);
    }
    GRE_HIP(hipStreamSynchronize(s));
  }

  // class starts
  {
    DevBuf cs;
    cs.alloc(8 * 9);
    hipLaunchKernelGGL(k_xclass, dim3(1), dim3(64), 0, s, key_s.as<uint16_t>(), E, per, cs.as<uint64_t>());
    GRE_HIP(hipGetLastError());
    GRE_HIP(hipMemcpyAsync(xp->class_start, cs.p, 8 * 9, hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
  }

  // segments: a new one wherever the key or the row changes
  DevBuf seg_row;
  {
    DevBuf flag, segid;
    flag.alloc(4 * E);
    segid.alloc(4 * E);
    xp->idx.alloc(4 * E);
    hipLaunchKernelGGL(k_xflags, dim3(gcap(E)), dim3(kB), 0, s, key_s.as<uint16_t>(), val_s.as<uint64_t>(), E,
                       flag.as<uint32_t>(), xp->idx.as<uint32_t>());
    GRE_HIP(hipGetLastError());
    size_t tb = 0;
    GRE_HIP(rocprim::inclusive_scan(nullptr, tb, flag.as<uint32_t>(), segid.as<uint32_t>(), (size_t)E,
                                    rocprim::plus<uint32_t>(), s));
    void* tmp = scratch(2, tb);
    GRE_HIP(rocprim::inclusive_scan(tmp, tb, flag.as<uint32_t>(), segid.as<uint32_t>(), (size_t)E,
                                    rocprim::plus<uint32_t>(), s));
    uint32_t ns = 0;
    GRE_HIP(hipMemcpyAsync(&ns, segid.as<uint32_t>() + (E - 1), 4, hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
    xp->nseg = ns;
    xp->seg_off.alloc(4 * ((size_t)ns + 1));
    seg_row.alloc(4 * (size_t)ns);
    hipLaunchKernelGGL(k_xsegs, dim3(gcap(E)), dim3(kB), 0, s, flag.as<uint32_t>(), segid.as<uint32_t>(),
                       val_s.as<uint64_t>(), E, xp->seg_off.as<uint32_t>(), seg_row.as<uint32_t>());
    GRE_HIP(hipGetLastError());
    const uint32_t eh = (uint32_t)E;
    GRE_HIP(hipMemcpyAsync(xp->seg_off.as<uint32_t>() + ns, &eh, 4, hipMemcpyHostToDevice, s));
    xp->K = 8 * per;
    xp->kseg.alloc(4 * (size_t)(xp->K + 1));
    hipLaunchKernelGGL(k_xkeyseg, dim3(1), dim3(kB), 0, s, key_s.as<uint16_t>(), segid.as<uint32_t>(), E, xp->K, ns,
                       xp->kseg.as<uint32_t>());
    GRE_HIP(hipGetLastError());
    GRE_HIP(hipStreamSynchronize(s));
  }
  key_s.release();
  val_s.release();

  // per-row segment lists, by range (stable sort of segment ids by row)
  const uint64_t nseg = xp->nseg;
  {
    DevBuf cnt, ids, rows_s;
    cnt.alloc(4 * ((size_t)nV + 1));
    GRE_HIP(hipMemsetAsync(cnt.p, 0, 4 * ((size_t)nV + 1), s));
    hipLaunchKernelGGL(k_xpcount, dim3(gcap(nseg)), dim3(kB), 0, s, seg_row.as<uint32_t>(), nseg, cnt.as<uint32_t>());
    GRE_HIP(hipGetLastError());
    xp->pair_off.alloc(4 * ((size_t)nV + 1));
    size_t tb = 0;
    GRE_HIP(rocprim::exclusive_scan(nullptr, tb, cnt.as<uint32_t>(), xp->pair_off.as<uint32_t>(), 0u,
                                    (size_t)nV + 1, rocprim::plus<uint32_t>(), s));
    void* tmp = scratch(2, tb);
    GRE_HIP(rocprim::exclusive_scan(tmp, tb, cnt.as<uint32_t>(), xp->pair_off.as<uint32_t>(), 0u, (size_t)nV + 1,
                                    rocprim::plus<uint32_t>(), s));
    ids.alloc(4 * nseg);
    rows_s.alloc(4 * nseg);
    xp->pair_list.alloc(4 * nseg);
    hipLaunchKernelGGL(k_iota32, dim3(gcap(nseg)), dim3(kB), 0, s, ids.as<uint32_t>(), nseg);
    GRE_HIP(hipGetLastError());
    tb = 0;
    const unsigned vb = bits_of(nV);
    GRE_HIP(rocprim::radix_sort_pairs(nullptr, tb, seg_row.as<uint32_t>(), rows_s.as<uint32_t>(),
                                      ids.as<uint32_t>(), xp->pair_list.as<uint32_t>(), (size_t)nseg, 0, vb, s));
    tmp = scratch(0, tb);
    GRE_HIP(rocprim::radix_sort_pairs(tmp, tb, seg_row.as<uint32_t>(), rows_s.as<uint32_t>(), ids.as<uint32_t>(),
                                      xp->pair_list.as<uint32_t>(), (size_t)nseg, 0, vb, s));
    GRE_HIP(hipStreamSynchronize(s));
  }
  g->xpart[mkey] = xp;
  return xp;
}

const XPart::PairTiles& xpart_pair_tiles(XPart& xp, uint32_t nV, uint32_t te, hipStream_t s) {
  auto it = xp.pair_tiles.find(te);
  if (it != xp.pair_tiles.end()) return it->second;
  XPart::PairTiles& T = xp.pair_tiles[te];
  uint64_t nt = (xp.nseg + te - 1) / te;
  if (nt == 0) nt = 1;
  T.ntiles = nt;
  T.rb.alloc(4 * (nt + 1));
  build_tile_rows(xp.pair_off.as<uint32_t>(), nV, nt, te, T.rb.as<uint32_t>(), s);
  // staged runs: LDS offsets are 16-bit, and a tile's keys fit its LDS table
  GRE_CHECK(te <= 65535 && xp.K > 0 && xp.K <= kMaxKeys, GRE_DEV_NULL, "pair tiles too large to stage");
  GRE_CHECK(te % 32 == 0, GRE_DEV_NULL, "pair tile entries must be a multiple of 32");
  const int K = xp.K;
  const uint64_t ns = xp.nseg;
  const uint32_t W = te / 32;
  DevBuf runA, runO;
  runA.alloc(4 * (nt + 1) * (uint64_t)K);
  runO.alloc(2 * nt * (uint64_t)(K + 1));
  T.words.alloc(8 * nt * (uint64_t)W);
  T.base.alloc(4 * nt * (uint64_t)K);
  GRE_HIP(hipMemsetAsync(T.words.p, 0, 8 * nt * (uint64_t)W, s));
  GRE_HIP(hipMemsetAsync(T.base.p, 0, 4 * nt * (uint64_t)K, s));
  // padded to whole tiles, so the row pass loads every slot of a tile unguarded
  T.slot.alloc(2 * nt * (uint64_t)te);
  GRE_HIP(hipMemsetAsync(T.slot.p, 0, 2 * nt * (uint64_t)te, s));
  {
    DevBuf pos;
    pos.alloc(4 * std::max<uint64_t>(ns, 1));
    hipLaunchKernelGGL(k_xpos, dim3(gcap(ns)), dim3(kB), 0, s, xp.pair_list.as<uint32_t>(), ns, pos.as<uint32_t>());
    GRE_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_xruna, dim3(gcap((nt + 1) * K)), dim3(kB), 0, s, xp.kseg.as<uint32_t>(), K,
                       pos.as<uint32_t>(), nt, te, runA.as<uint32_t>());
    GRE_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_xruno, dim3(gcap(nt)), dim3(kB), 0, s, runA.as<uint32_t>(), K, nt, runO.as<uint16_t>());
    GRE_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_xslot, dim3(gcap(ns)), dim3(kB), 0, s, xp.pair_list.as<uint32_t>(), ns, te,
                       xp.kseg.as<uint32_t>(), K, runA.as<uint32_t>(), runO.as<uint16_t>(),
                       T.slot.as<uint16_t>());
    GRE_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_xstage, dim3(gcap(nt)), dim3(kB), 0, s, runA.as<uint32_t>(), runO.as<uint16_t>(), K, nt, W,
                       T.words.as<uint2>(), T.base.as<uint32_t>());
    GRE_HIP(hipGetLastError());
    GRE_HIP(hipStreamSynchronize(s));
  }
  return T;
}

const XPart::Tiles& xpart_tiles(XPart& xp, uint32_t te, hipStream_t s) {
  auto it = xp.tiles.find(te);
  if (it != xp.tiles.end()) return it->second;
  XPart::Tiles& T = xp.tiles[te];
  const uint64_t E = xp.class_start[8];
  uint64_t nt = (E + te - 1) / te;
  if (nt == 0) nt = 1;
  T.ntiles = nt;
  T.rb.alloc(4 * (nt + 1));
  build_tile_rows(xp.seg_off.as<uint32_t>(), (uint32_t)xp.nseg, nt, te, T.rb.as<uint32_t>(), s);
  // class c runs tiles [xm[c], xm[c+1]) on the workgroups with id % 8 == c
  uint32_t xm[9];
  for (int c = 0; c < 8; ++c) xm[c] = (uint32_t)std::min<uint64_t>(xp.class_start[c] / te, nt);
  xm[0] = 0;
  xm[8] = (uint32_t)nt;
  uint32_t most = 0;
  for (int c = 0; c < 8; ++c) most = std::max(most, xm[c + 1] - xm[c]);
  T.grid = 8u * std::max(most, 1u);
  T.xmap.alloc(4 * 9);
  GRE_HIP(hipMemcpyAsync(T.xmap.p, xm, 4 * 9, hipMemcpyHostToDevice, s));
  GRE_HIP(hipStreamSynchronize(s));
  return T;
}

}  // namespace gre
