// Traversals outside GRE's vertex-program model, used as the plain
// single-source path and the batched multi-source path:
//   gre_bfs    direction-optimising BFS (levels): top-down push over CSR with
//              an edge-balanced frontier, bottom-up pull over CSC with early
//              exit, switched by frontier/unvisited edge counts;
//   gre_msbfs  up to 64 BFS sources at once, one 64-bit mask per vertex,
//              direction-optimising like gre_bfs: top-down push of the
//              frontier masks with 64-bit atomic ORs, bottom-up pull over CSC
//              with early exit once a vertex holds every source bit;
//   gre_bc     Brandes BC, FP64 path counts and dependencies, 64 sources per
//              chunk as the 64 lanes of a wave, pull over CSC (forward) and
//              CSR (backward), deterministic sums.
#include "graph.hpp"

#include <algorithm>
#include <map>
#include <type_traits>
#include <vector>

namespace gre {
namespace {

constexpr int kB = 256;
constexpr int kEB = kB * 8;

unsigned gcap(uint64_t n) {
  uint64_t g = (n + kB - 1) / kB;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

template <typename T>
__global__ void k_fill(T* p, uint64_t n, T v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---------------------------------------------------------------------------
// direction-optimising BFS
// ---------------------------------------------------------------------------
__global__ void k_bfs_seed(int32_t* level, uint32_t src, uint32_t* queue, uint32_t* fbits) {
  level[src] = 0;
  queue[0] = src;
  fbits[src >> 5] |= 1u << (src & 31);
}

__global__ void k_qdeg(const uint32_t* __restrict__ q, uint32_t n, const uint64_t* __restrict__ off,
                       uint64_t* __restrict__ deg) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x)
    deg[i] = i < n ? off[q[i] + 1] - off[q[i]] : 0;
}

// top-down: edge-balanced expansion of the queue; new vertices claim level L+1
__global__ __launch_bounds__(kB) void k_td(const uint64_t* __restrict__ off, const uint32_t* __restrict__ idx,
                                           const uint32_t* __restrict__ q, const uint64_t* __restrict__ qpos,
                                           uint32_t qn, uint64_t T, int32_t L, int32_t* __restrict__ level,
                                           uint32_t* __restrict__ nq, uint32_t* __restrict__ ncount,
                                           uint32_t* __restrict__ nbits, unsigned long long* __restrict__ ndeg) {
  __shared__ uint32_t range[2];
  const uint64_t k0 = (uint64_t)blockIdx.x * kEB;
  const uint64_t k1 = k0 + kEB < T ? k0 + kEB : T;
  if (threadIdx.x < 2) {
    const uint64_t key = threadIdx.x == 0 ? k0 : k1 - 1;
    uint32_t lo = 0, hi = qn;
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (qpos[mid] <= key) lo = mid; else hi = mid;
    }
    range[threadIdx.x] = lo;
  }
  __syncthreads();
  const uint32_t i0 = range[0], i1 = range[1];
  uint64_t mydeg = 0;
  // a fixed number of rounds for every thread keeps the block-level append uniform
  for (int round = 0; round < kEB / kB; ++round) {
    const uint64_t k = k0 + (uint64_t)round * kB + threadIdx.x;
    bool take = false;
    uint32_t v = 0;
    if (k < k1) {
      uint32_t lo = i0, hi = i1 + 1;
      while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (qpos[mid] <= k) lo = mid; else hi = mid;
      }
      const uint32_t u = q[lo];
      v = idx[off[u] + (k - qpos[lo])];
      if (level[v] < 0 && atomicCAS(level + v, -1, L + 1) == -1) {
        take = true;
        atomicOr(nbits + (v >> 5), 1u << (v & 31));
        mydeg += off[v + 1] - off[v];
      }
    }
    block_append(take, v, nq, ncount);
  }
  block_add(static_cast<unsigned long long>(mydeg), ndeg);
}

// bottom-up: every unvisited vertex looks for a parent in the frontier bitmap
__global__ __launch_bounds__(kB) void k_bu(const uint64_t* __restrict__ coff, const uint32_t* __restrict__ cidx,
                                           const uint64_t* __restrict__ off, uint32_t nV, int32_t L,
                                           const uint32_t* __restrict__ fbits, int32_t* __restrict__ level,
                                           uint32_t* __restrict__ nbits, uint32_t* __restrict__ ncount,
                                           unsigned long long* __restrict__ ndeg) {
  uint32_t cnt = 0;
  uint64_t dsum = 0;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV; v += (uint64_t)gridDim.x * blockDim.x) {
    if (level[v] >= 0) continue;
    const uint64_t b = coff[v], e = coff[v + 1];
    // 8 in-neighbours per step, their frontier bits read together (past the
    // row's end: the first neighbour again, which changes nothing)
    bool found = false;
    for (uint64_t p = b; p < e && !found; p += 8) {
      uint32_t w[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) w[j] = cidx[p + j < e ? p + j : b];
      uint32_t fb[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) fb[j] = fbits[w[j] >> 5];
#pragma unroll
      for (int j = 0; j < 8; ++j) found = found || ((fb[j] >> (w[j] & 31)) & 1u);
    }
    if (found) {
      level[v] = L + 1;
      atomicOr(nbits + (v >> 5), 1u << (v & 31));
      ++cnt;
      dsum += off[v + 1] - off[v];
    }
  }
  block_add2(cnt, ncount, static_cast<unsigned long long>(dsum), ndeg);
}

// bitmap -> queue (order within the queue does not matter to BFS levels)
__global__ void k_bits_to_queue(const uint32_t* __restrict__ bits, uint32_t nwords, uint32_t* __restrict__ q,
                                uint32_t* __restrict__ qn) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // every thread runs the same number of rounds (block-level reservation)
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < nwords; base += stride) {
    const uint64_t w = base + threadIdx.x;
    uint32_t m = w < nwords ? bits[w] : 0u;
    const uint32_t p = block_reserve((uint32_t)__popc(m), qn);
    uint32_t i = 0;
    while (m) {
      const int b = __ffs(m) - 1;
      q[p + i++] = (uint32_t)(w * 32 + b);
      m &= m - 1;
    }
  }
}

struct BfsBufs {
  DevBuf q[2], qdeg, qpos, bits[2], cnt, deg, lev;
  uint32_t nV = 0;
};
BfsBufs& bfs_bufs(uint32_t nV) {
  static BfsBufs& b = *new BfsBufs();  // lives until the process ends
  if (b.nV < nV || !b.q[0].p) {
    b.q[0].alloc(4 * (size_t)nV);
    b.q[1].alloc(4 * (size_t)nV);
    b.qpos.alloc(8 * ((size_t)nV + 1));
    b.qdeg.alloc(8 * ((size_t)nV + 1));
    const size_t nw = ((size_t)nV + 63) / 64 * 2;
    b.bits[0].alloc(4 * nw);
    b.bits[1].alloc(4 * nw);
    b.cnt.alloc(64);
    b.deg.alloc(64);
    b.lev.alloc(4 * (size_t)nV);   // levels when the caller keeps none on the device
    b.nV = nV;
  }
  return b;
}

void bfs_run(gre_graph* g, uint32_t src, int32_t* level, uint64_t* edges_out, int32_t* depth_out) {
  hipStream_t s = stream();
  const uint32_t nV = g->nV;
  BfsBufs& b = bfs_bufs(nV);
  const uint32_t nwords = (uint32_t)(((uint64_t)nV + 63) / 64 * 2);
  const uint64_t* off = g->off.as<uint64_t>();
  const uint64_t* coff = g->coff.as<uint64_t>();
  hipLaunchKernelGGL(k_fill<int32_t>, dim3(gcap(nV)), dim3(kB), 0, s, level, (uint64_t)nV, -1);
  GRE_HIP(hipMemsetAsync(b.bits[0].p, 0, 4 * (size_t)nwords, s));
  hipLaunchKernelGGL(k_bfs_seed, dim3(1), dim3(1), 0, s, level, src, b.q[0].as<uint32_t>(), b.bits[0].as<uint32_t>());
  GRE_HIP(hipGetLastError());
  uint64_t dsrc = 0;
  {
    uint64_t o2[2];
    GRE_HIP(hipMemcpyAsync(o2, off + src, 16, hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
    dsrc = o2[1] - o2[0];
  }
  uint32_t qn = 1;
  uint64_t reached_deg = dsrc;        // out-degree sum of the vertices reached so far
  uint64_t mf = dsrc;                 // edges out of the frontier
  uint64_t mu = g->nE - dsrc;         // edges of unvisited vertices (approx.)
  int cur = 0;
  bool topdown = true;
  int32_t L = 0;
  bool queue_valid = true;            // q[cur] holds the frontier
  const double alpha = 40.0, beta = 24.0;
  uint32_t prev_n = 0;
  while (qn > 0) {
    const int nxt = cur ^ 1;
    if (topdown && (double)mf > (double)mu / alpha) topdown = false;
    else if (!topdown && (double)qn < (double)nV / beta && qn < prev_n) topdown = true;
    GRE_HIP(hipMemsetAsync(b.bits[nxt].p, 0, 4 * (size_t)nwords, s));
    GRE_HIP(hipMemsetAsync(b.cnt.p, 0, 8, s));
    GRE_HIP(hipMemsetAsync(b.deg.p, 0, 8, s));
    if (topdown) {
      if (!queue_valid) {
        GRE_HIP(hipMemsetAsync(b.cnt.p, 0, 4, s));
        hipLaunchKernelGGL(k_bits_to_queue, dim3(gcap(nwords)), dim3(kB), 0, s, b.bits[cur].as<uint32_t>(), nwords,
                           b.q[cur].as<uint32_t>(), b.cnt.as<uint32_t>());
        GRE_HIP(hipMemsetAsync(b.cnt.p, 0, 4, s));
      }
      hipLaunchKernelGGL(k_qdeg, dim3(gcap((uint64_t)qn + 1)), dim3(kB), 0, s, b.q[cur].as<uint32_t>(), qn, off,
                         b.qdeg.as<uint64_t>());
      scan_excl_u64(b.qdeg.as<uint64_t>(), b.qpos.as<uint64_t>(), (size_t)qn + 1, s);
      uint64_t T = 0;
      GRE_HIP(hipMemcpyAsync(&T, b.qpos.as<uint64_t>() + qn, 8, hipMemcpyDeviceToHost, s));
      GRE_HIP(hipStreamSynchronize(s));
      if (T > 0) {
        ProfScope ps("bfs_topdown");
        hipLaunchKernelGGL(k_td, dim3((unsigned)((T + kEB - 1) / kEB)), dim3(kB), 0, s, off, g->idx.as<uint32_t>(),
                           b.q[cur].as<uint32_t>(), b.qpos.as<uint64_t>(), qn, T, L, level, b.q[nxt].as<uint32_t>(),
                           b.cnt.as<uint32_t>(), b.bits[nxt].as<uint32_t>(), (unsigned long long*)b.deg.p);
      }
      queue_valid = true;
    } else {
      ProfScope ps("bfs_bottomup");
      hipLaunchKernelGGL(k_bu, dim3(gcap(nV)), dim3(kB), 0, s, coff, g->cidx.as<uint32_t>(), off, nV, L,
                         b.bits[cur].as<uint32_t>(), level, b.bits[nxt].as<uint32_t>(), b.cnt.as<uint32_t>(),
                         (unsigned long long*)b.deg.p);
      queue_valid = false;
    }
    GRE_HIP(hipGetLastError());
    uint32_t nn = 0;
    uint64_t nd = 0;
    GRE_HIP(hipMemcpyAsync(&nn, b.cnt.p, 4, hipMemcpyDeviceToHost, s));
    GRE_HIP(hipMemcpyAsync(&nd, b.deg.p, 8, hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
    prev_n = qn;
    qn = nn;
    mf = nd;
    reached_deg += nd;
    mu = mu > nd ? mu - nd : 0;
    cur = nxt;
    ++L;
  }
  // every reached vertex entered exactly one frontier, whose out-degree sum
  // the level reported: the reached edges are their total; the deepest level
  // is the last one that found vertices
  if (edges_out) *edges_out = reached_deg;
  if (depth_out) *depth_out = L - 1;
}

// ---------------------------------------------------------------------------
// multi-source BFS: 64 sources as bits of one mask per vertex
// ---------------------------------------------------------------------------
// Working arrays, grown to the largest graph seen and kept for the process
// (like the single-source ones): a call allocates nothing of size V.
// Level control of the 64-source BFS, in device memory: every level kernel
// takes its sizes, its direction and the level number from here, and two
// one-thread kernels advance it, so levels run back to back with no host
// round trip (a captured graph of kMsBatch levels is replayed until mode is
// kMsDone).
enum : uint32_t { kMsTopDown = 0, kMsBottomUp = 1, kMsDone = 2 };
struct MsCtl {
  uint32_t L;          // level being expanded (frontier = vertices first reached at L)
  uint32_t qn;         // frontier size
  uint32_t nn;         // next frontier size (appended to by the level's kernels)
  uint32_t prev_n;     // the frontier size of the level before
  uint32_t mode;
  uint32_t pad;
  uint64_t T;          // out-edges of the frontier (top-down levels)
  uint64_t mf;         // out-edges of the frontier, from the level before
  uint64_t mu;         // edges of the vertices still missing source bits
  unsigned long long sums[2];   // the level's (new frontier's edges, edges of vertices now complete)
};
constexpr int kMsBatch = 8;     // levels per replayed graph (even: queue parity repeats)
constexpr int kScanB = 2048;    // elements per chunk of the device-length scan
// Grids of the level kernels: grid-stride loops over their work, sized so a
// level's kernels for the direction not taken (which return at once) cost
// little to dispatch, while a full one still has many more waves than the
// chip holds.
constexpr unsigned kMsGrid = 8192;
constexpr unsigned kMsScanGrid = 1024;

// Control of the batched Brandes BC, in device memory.  A step is one forward
// level, one backward level, or a chunk's set-up fused with its first forward
// level; the control says which, every kernel of a step reads it and returns
// at once when the step is not its kind, so one launch sequence of kBcBatch
// steps is captured and replayed over every 64-source chunk until the
// control reports done.
enum : uint32_t { kBcInit = 0, kBcFwd = 1, kBcBwd = 2, kBcDone = 3, kBcTooDeep = 4 };
constexpr uint32_t kBcMaxDepth = 250;  // levels are kept as uint8 (0xff: not reached)
constexpr int kBcBatch = 16;           // steps per replayed graph
struct BcCtl {
  uint32_t phase;
  uint32_t mode;       // forward direction: kMsTopDown or kMsBottomUp
  uint32_t chunk;      // chunk being run: sources [64 chunk, 64 chunk + nc)
  uint32_t nchunks;
  uint32_t nsrc;
  uint32_t nc;
  uint32_t L;          // forward: level being expanded; backward: level being folded
  uint32_t depth;      // deepest level of the chunk's forward pass
  uint32_t qn;         // frontier size (the frontier is q[L & 1], its lane masks fr[L & 1])
  uint32_t nn;         // next frontier size (appended to by the level's kernels)
  uint64_t full;       // lane mask of the chunk's sources
  uint64_t T;          // top-down: out-edges of the frontier (prefix total)
  uint64_t mf;         // out-edges of the frontier
  unsigned long long sums[2];   // sums[0]: out-edges of the next frontier
};
__device__ inline bool td_on(const MsCtl* c) { return c->mode == kMsTopDown; }
__device__ inline bool td_on(const BcCtl* c) { return c->phase == kBcFwd && c->mode == kMsTopDown; }
__device__ inline bool bu_on(const BcCtl* c) { return c->phase == kBcFwd && c->mode == kMsBottomUp; }

struct MsBufs {
  DevBuf vis, fr[2], nxt, q[2], qdeg, qpos, flag, sums, src, ctl, bsum;
  uint32_t nV = 0;
  MsCtl* ctl_host = nullptr;    // pinned
};
MsBufs& ms_bufs(uint32_t nV) {
  static MsBufs& b = *new MsBufs();  // lives until the process ends
  if (b.nV < nV || !b.vis.p) {
    b.vis.alloc(8 * (size_t)nV);
    b.fr[0].alloc(8 * (size_t)nV);
    b.fr[1].alloc(8 * (size_t)nV);
    b.nxt.alloc(8 * (size_t)nV);
    b.q[0].alloc(4 * (size_t)nV);
    b.q[1].alloc(4 * (size_t)nV);
    b.qdeg.alloc(8 * ((size_t)nV + 1));
    b.qpos.alloc(8 * ((size_t)nV + 1));
    b.flag.alloc(16);
    b.sums.alloc(16);
    b.src.alloc(4 * 64);
    b.ctl.alloc(sizeof(MsCtl));
    b.bsum.alloc(8 * (((size_t)nV + 1 + kScanB - 1) / kScanB + 1));
    if (!b.ctl_host) GRE_HIP(hipHostMalloc((void**)&b.ctl_host, sizeof(MsCtl), hipHostMallocDefault));
    b.nV = nV;
  }
  return b;
}

// the direction of the level, as the host loop chose it before: bottom-up once
// the frontier's edges exceed 1/14 of the edges still missing bits, top-down
// again once the frontier shrinks under V/24; done once it is empty
__global__ void k_ms_ctl_begin(MsCtl* c, uint32_t nV) {
  if (c->mode == kMsDone) return;
  if (c->qn == 0) { c->mode = kMsDone; return; }
  bool td = c->mode == kMsTopDown;
  if (td && c->L > 0 && c->mf * 14ull > c->mu) td = false;
  else if (!td && (uint64_t)c->qn * 24ull < (uint64_t)nV && c->qn < c->prev_n) td = true;
  c->mode = td ? kMsTopDown : kMsBottomUp;
  c->nn = 0;
  c->T = 0;
  c->sums[0] = 0;
  c->sums[1] = 0;
}

__global__ void k_ms_ctl_end(MsCtl* c) {
  if (c->mode == kMsDone) return;
  c->prev_n = c->qn;
  c->qn = c->nn;
  c->mf = c->sums[0];
  c->mu = c->mu > c->sums[1] ? c->mu - c->sums[1] : 0;
  c->L += 1;
}

// top-down levels: out-degrees of the frontier, then their exclusive prefix
// (T = the total) by a three-pass scan whose length is read from the control
__global__ void k_ms_qdeg(const MsCtl* __restrict__ c, const uint32_t* __restrict__ q,
                          const uint64_t* __restrict__ off, uint64_t* __restrict__ d) {
  if (!td_on(c)) return;
  const uint32_t n = c->qn;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x)
    d[i] = i < n ? off[q[i] + 1] - off[q[i]] : 0;
}

template <class Ctl>
__global__ __launch_bounds__(kB) void k_ms_scan_sums(const Ctl* __restrict__ c, const uint64_t* __restrict__ in,
                                                     uint64_t* __restrict__ bsum) {
  __shared__ uint64_t red[kB / 64];
  if (!td_on(c)) return;
  const uint64_t n = (uint64_t)c->qn + 1;
  for (uint64_t b = blockIdx.x; b * kScanB < n; b += gridDim.x) {
    const uint64_t b0 = b * kScanB;
    uint64_t t = 0;
    for (int k = 0; k < kScanB / kB; ++k) {
      const uint64_t i = b0 + (uint64_t)threadIdx.x * (kScanB / kB) + k;
      if (i < n) t += in[i];
    }
    t = wave_sum(t);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) bsum[b] = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
  }
}

template <class Ctl>
__global__ __launch_bounds__(1024) void k_ms_scan_top(const Ctl* __restrict__ c, uint64_t* __restrict__ bsum) {
  __shared__ uint64_t wt[16];
  if (!td_on(c)) return;
  const uint64_t nb = ((uint64_t)c->qn + 1 + kScanB - 1) / kScanB;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t carry = 0;
  for (uint64_t base = 0; base < nb; base += 1024) {
    const uint64_t i = base + threadIdx.x;
    const uint64_t v = i < nb ? bsum[i] : 0;
    uint64_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
      const uint64_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wt[w] = x;
    __syncthreads();
    uint64_t pre = carry;
    for (int k = 0; k < w; ++k) pre += wt[k];
    if (i < nb) bsum[i] = pre + x - v;
    for (int k = 0; k < 16; ++k) carry += wt[k];
    __syncthreads();
  }
}

template <class Ctl>
__global__ __launch_bounds__(kB) void k_ms_scan_apply(Ctl* __restrict__ c, const uint64_t* __restrict__ in,
                                                      const uint64_t* __restrict__ bsum, uint64_t* __restrict__ out) {
  __shared__ uint64_t wt[kB / 64];
  if (!td_on(c)) return;
  const uint64_t n = (uint64_t)c->qn + 1;
  for (uint64_t b = blockIdx.x; b * kScanB < n; b += gridDim.x) {
    const uint64_t b0 = b * kScanB;
    constexpr int P = kScanB / kB;
    const uint64_t i0 = b0 + (uint64_t)threadIdx.x * P;
    uint64_t v[P];
    uint64_t t = 0;
#pragma unroll
    for (int k = 0; k < P; ++k) {
      v[k] = i0 + k < n ? in[i0 + k] : 0;
      t += v[k];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t x = t;
    for (int o = 1; o < 64; o <<= 1) {
      const uint64_t y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wt[w] = x;
    __syncthreads();
    uint64_t r = bsum[b] + x - t;
    for (int k = 0; k < w; ++k) r += wt[k];
#pragma unroll
    for (int k = 0; k < P; ++k) {
      if (i0 + k < n) {
        out[i0 + k] = r;
        if (i0 + k == n - 1) c->T = r;   // qdeg[qn] = 0: the prefix there is the total
      }
      r += v[k];
    }
    __syncthreads();   // wt is rewritten by the next chunk
  }
}

__global__ void k_ms_seed(const uint32_t* __restrict__ src, int n, unsigned long long* vis, unsigned long long* fr,
                          int32_t* lev, uint32_t nV) {
  for (int i = 0; i < n; ++i) {
    vis[src[i]] |= 1ULL << i;
    fr[src[i]] |= 1ULL << i;
    if (lev) lev[(uint64_t)i * nV + src[i]] = 0;
  }
}

// Records level L+1 for every source bit in m at vertex v.
__device__ inline void ms_record(int32_t* __restrict__ lev, uint32_t nV, uint32_t v, unsigned long long m,
                                 int32_t L) {
  while (m) {
    const int i = __ffsll(m) - 1;
    lev[(uint64_t)i * nV + v] = L + 1;
    m &= m - 1;
  }
}

// top-down: the frontier's out-edges, edge-balanced; a vertex that gains new
// source bits is queued by the first edge that reaches it (nxt was zero)
__global__ __launch_bounds__(kB) void k_ms_td(const uint64_t* __restrict__ off, const uint32_t* __restrict__ idx,
                                              const uint32_t* __restrict__ q, const uint64_t* __restrict__ qpos,
                                              MsCtl* __restrict__ c, const unsigned long long* __restrict__ fr,
                                              const unsigned long long* __restrict__ vis,
                                              unsigned long long* __restrict__ nxt, uint32_t* __restrict__ nq) {
  __shared__ uint32_t range[2];
  if (c->mode != kMsTopDown) return;
  const uint32_t qn = c->qn;
  const uint64_t T = c->T;
  uint32_t* ncount = &c->nn;
  // edge blocks of kEB, grid-stride (the same trip count on every thread of a block)
  for (uint64_t blk = blockIdx.x; blk * kEB < T; blk += gridDim.x) {
    const uint64_t k0 = blk * kEB;
    const uint64_t k1 = k0 + kEB < T ? k0 + kEB : T;
    if (threadIdx.x < 2) {
      const uint64_t key = threadIdx.x == 0 ? k0 : k1 - 1;
      uint32_t lo = 0, hi = qn;
      while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (qpos[mid] <= key) lo = mid; else hi = mid;
      }
      range[threadIdx.x] = lo;
    }
    __syncthreads();
    const uint32_t i0 = range[0], i1 = range[1];
    // a fixed number of rounds for every lane keeps the wave-level append uniform
    for (int round = 0; round < kEB / kB; ++round) {
      const uint64_t k = k0 + (uint64_t)round * kB + threadIdx.x;
      bool take = false;
      uint32_t v = 0;
      if (k < k1) {
        uint32_t lo = i0, hi = i1 + 1;
        while (hi - lo > 1) {
          const uint32_t mid = lo + (hi - lo) / 2;
          if (qpos[mid] <= k) lo = mid; else hi = mid;
        }
        const uint32_t u = q[lo];
        v = idx[off[u] + (k - qpos[lo])];
        const unsigned long long bits = fr[u] & ~vis[v];
        if (bits) take = atomicOr(nxt + v, bits) == 0ULL;
      }
      block_append(take, v, nq, ncount);
    }
    __syncthreads();   // range[] is rewritten by the next edge block
  }
}

// top-down, second pass over the queued vertices: publish the new bits
__global__ void k_ms_td_fin(const uint32_t* __restrict__ nq, MsCtl* __restrict__ c, unsigned long long full,
                            unsigned long long* __restrict__ nxt, unsigned long long* __restrict__ nfr,
                            unsigned long long* __restrict__ vis, const uint64_t* __restrict__ off,
                            int32_t* __restrict__ lev, uint32_t nV) {
  if (c->mode != kMsTopDown) return;
  const uint32_t nn = c->nn;
  const int32_t L = (int32_t)c->L;
  unsigned long long* sums = c->sums;
  uint64_t deg = 0, fulldeg = 0;
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < nn; j += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = nq[j];
    const unsigned long long seen = vis[v];
    const unsigned long long m = nxt[v] & ~seen;
    nxt[v] = 0ULL;
    nfr[v] = m;
    vis[v] = seen | m;
    const uint64_t d = off[v + 1] - off[v];
    deg += d;
    if ((seen | m) == full) fulldeg += d;
    if (lev) ms_record(lev, nV, v, m, L);
  }
  block_add2(static_cast<unsigned long long>(deg), sums, static_cast<unsigned long long>(fulldeg), sums + 1);
}

// Rows with more than `lng` in-edges are handled apart in the bottom-up step:
// one thread would walk a hub's million in-edges alone, and a vertex rarely
// gains all 64 bits in one level, so early levels scan whole rows.
// The threshold and the chunk length (in-edges per chunk, one wave each) are
// picked per graph by timing the candidates below on the first call.
constexpr uint64_t kMsCand[][2] = {{1024, 4096}, {256, 1024}, {4096, 16384}, {512, 2048}};
// calls with fewer sources run the first pair and leave the pick to a later call
constexpr int kTuneMinSources = 32;

// bottom-up: every vertex still missing source bits ORs its in-neighbours'
// frontier masks, stopping once it has them all (rows above lng in-edges:
// k_ms_bu_chunks and k_ms_bu_long_fin)
__global__ __launch_bounds__(kB) void k_ms_bu(const uint64_t* __restrict__ coff, const uint32_t* __restrict__ cidx,
                                              const uint64_t* __restrict__ off, uint32_t nV, unsigned long long full,
                                              const unsigned long long* __restrict__ fr,
                                              unsigned long long* __restrict__ nfr,
                                              unsigned long long* __restrict__ vis, int32_t* __restrict__ lev,
                                              MsCtl* __restrict__ c, uint32_t* __restrict__ nq, uint64_t lng) {
  if (c->mode != kMsBottomUp) return;
  const int32_t L = (int32_t)c->L;
  uint32_t* ncount = &c->nn;
  unsigned long long* sums = c->sums;
  uint64_t deg = 0, fulldeg = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // every lane runs the same number of rounds so the wave-level append is uniform
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < nV; base += stride) {
    const uint64_t v = base + threadIdx.x;
    bool take = false;
    if (v < nV) {
      const unsigned long long seen = vis[v];
      const uint64_t b = coff[v], e = coff[v + 1];
      if (seen != full && e - b <= lng) {
        unsigned long long m = 0;
        // 8 in-neighbours per step, their masks gathered together (past the
        // row's end: the first neighbour again); once (m | seen) == full the new
        // bits m & ~seen are all the missing ones, whichever neighbours were ORed
        for (uint64_t p = b; p < e; p += 8) {
          uint32_t w[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) w[j] = cidx[p + j < e ? p + j : b];
          unsigned long long f[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) f[j] = fr[w[j]];
#pragma unroll
          for (int j = 0; j < 8; ++j) m |= f[j];
          if ((m | seen) == full) break;
        }
        m &= ~seen;
        if (m) {
          nfr[v] = m;
          vis[v] = seen | m;
          take = true;
          const uint64_t d = off[v + 1] - off[v];
          deg += d;
          if ((seen | m) == full) fulldeg += d;
          if (lev) ms_record(lev, nV, (uint32_t)v, m, L);
        }
      }
    }
    block_append(take, (uint32_t)v, nq, ncount);
  }
  block_add2(static_cast<unsigned long long>(deg), sums, static_cast<unsigned long long>(fulldeg), sums + 1);
}

// bottom-up over the long rows' chunks, one wave per chunk: the OR of the
// chunk's in-neighbours' frontier masks goes to acc[row] (stopping once the
// row has every bit)
__global__ __launch_bounds__(kB) void k_ms_bu_chunks(const uint64_t* __restrict__ coff,
                                                     const uint32_t* __restrict__ cidx,
                                                     const uint32_t* __restrict__ rows,
                                                     const uint32_t* __restrict__ cv,
                                                     const uint64_t* __restrict__ cb, uint32_t nchunks,
                                                     unsigned long long full,
                                                     const unsigned long long* __restrict__ fr,
                                                     const unsigned long long* __restrict__ vis,
                                                     unsigned long long* __restrict__ acc, const MsCtl* __restrict__ c,
                                                     uint64_t chunk) {
  if (c->mode != kMsBottomUp) return;
  const int lane = (int)(threadIdx.x & 63);
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t c = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; c < nchunks; c += nw) {
    const uint32_t i = cv[c];
    const uint32_t v = rows[i];
    const unsigned long long seen = vis[v];
    if (seen == full) continue;  // the same chunk, so the same answer, on every lane
    const uint64_t b = cb[c], rowend = coff[v + 1];
    const uint64_t e = b + chunk < rowend ? b + chunk : rowend;
    unsigned long long m = 0;
    for (uint64_t p = b; p < e; p += 4 * 64) {
      // 4 in-neighbours per lane per step, gathered together (past the end:
      // the chunk's first in-neighbour again, which changes nothing)
      uint32_t w[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint64_t q = p + (uint64_t)(j * 64 + lane);
        w[j] = cidx[q < e ? q : b];
      }
      unsigned long long f[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) f[j] = fr[w[j]];
#pragma unroll
      for (int j = 0; j < 4; ++j) m |= f[j];
      if ((wave_or(m) | seen) == full) break;
    }
    m = wave_or(m) & ~seen;
    if (lane == 0 && m) atomicOr(acc + i, m);
  }
}

// the long rows' new bits (acc, cleared for the next level), as k_ms_bu does
// for the others
__global__ __launch_bounds__(kB) void k_ms_bu_long_fin(const uint32_t* __restrict__ rows, uint32_t n,
                                                       unsigned long long full,
                                                       unsigned long long* __restrict__ acc,
                                                       unsigned long long* __restrict__ nfr,
                                                       unsigned long long* __restrict__ vis,
                                                       const uint64_t* __restrict__ off, int32_t* __restrict__ lev,
                                                       uint32_t nV, MsCtl* __restrict__ c, uint32_t* __restrict__ nq) {
  if (c->mode != kMsBottomUp) return;
  const int32_t L = (int32_t)c->L;
  uint32_t* ncount = &c->nn;
  unsigned long long* sums = c->sums;
  uint64_t deg = 0, fulldeg = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const uint64_t i = base + threadIdx.x;
    bool take = false;
    uint32_t v = 0;
    if (i < n) {
      v = rows[i];
      const unsigned long long seen = vis[v];
      const unsigned long long m = acc[i] & ~seen;
      acc[i] = 0ULL;
      if (m) {
        nfr[v] = m;
        vis[v] = seen | m;
        take = true;
        const uint64_t d = off[v + 1] - off[v];
        deg += d;
        if ((seen | m) == full) fulldeg += d;
        if (lev) ms_record(lev, nV, v, m, L);
      }
    }
    block_append(take, v, nq, ncount);
  }
  block_add2(static_cast<unsigned long long>(deg), sums, static_cast<unsigned long long>(fulldeg), sums + 1);
}

// vertices whose row in off has more than thr entries
__global__ void k_long_rows(const uint64_t* __restrict__ off, uint32_t nV, uint64_t thr, uint32_t* __restrict__ rows,
                            uint32_t* __restrict__ n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < nV; base += stride) {
    const uint64_t v = base + threadIdx.x;
    const bool take = v < nV && off[v + 1] - off[v] > thr;
    block_append(take, (uint32_t)v, rows, n);
  }
}

__global__ void k_row_bounds_of(const uint32_t* __restrict__ rows, uint32_t n, const uint64_t* __restrict__ off,
                                uint64_t* __restrict__ be) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    be[2 * i] = off[rows[i]];
    be[2 * i + 1] = off[rows[i] + 1];
  }
}

// the rows of the CSC (csc) or CSR with more than thr entries and their
// chunks of up to chunk entries, built once per graph and (csc, thr, chunk)
LongRows& long_rows(gre_graph* g, bool csc, uint64_t thr, uint64_t chunk, hipStream_t s) {
  const uint64_t key = ((uint64_t)csc << 63) | (thr << 32) | chunk;
  auto it = g->long_rows.find(key);
  if (it != g->long_rows.end()) return it->second;
  const uint32_t nV = g->nV;
  const uint64_t* off = (csc ? g->coff : g->off).as<uint64_t>();
  DevBuf rows, cnt;
  rows.alloc(4 * (size_t)nV);
  cnt.alloc(4);
  GRE_HIP(hipMemsetAsync(cnt.p, 0, 4, s));
  hipLaunchKernelGGL(k_long_rows, dim3(gcap(nV)), dim3(kB), 0, s, off, nV, thr, rows.as<uint32_t>(),
                     cnt.as<uint32_t>());
  GRE_HIP(hipGetLastError());
  uint32_t n = 0;
  GRE_HIP(hipMemcpyAsync(&n, cnt.p, 4, hipMemcpyDeviceToHost, s));
  GRE_HIP(hipStreamSynchronize(s));
  std::vector<uint32_t> hr(n);
  std::vector<uint64_t> hbe(2 * (size_t)n);
  if (n) {
    DevBuf be;
    be.alloc(16 * (size_t)n);
    hipLaunchKernelGGL(k_row_bounds_of, dim3(gcap(n)), dim3(kB), 0, s, rows.as<uint32_t>(), n, off,
                       be.as<uint64_t>());
    GRE_HIP(hipGetLastError());
    GRE_HIP(hipMemcpyAsync(hr.data(), rows.p, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
    GRE_HIP(hipMemcpyAsync(hbe.data(), be.p, 16 * (size_t)n, hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
  }
  std::vector<uint32_t> cv, cs;
  std::vector<uint64_t> cb;
  for (uint32_t i = 0; i < n; ++i) {
    cs.push_back((uint32_t)cv.size());
    for (uint64_t p = hbe[2 * (size_t)i]; p < hbe[2 * (size_t)i + 1]; p += chunk) {
      cv.push_back(i);
      cb.push_back(p);
    }
  }
  GRE_CHECK(cv.size() < 0xffffffffULL, GRE_DEV_FAILURE, "too many long-row chunks");
  cs.push_back((uint32_t)cv.size());
  LongRows& lr = g->long_rows[key];
  lr.n = n;
  lr.nchunks = (uint32_t)cv.size();
  lr.rows.alloc(4 * std::max<size_t>(n, 1));
  lr.acc.alloc(8 * std::max<size_t>(n, 1));
  lr.cv.alloc(4 * std::max<size_t>(cv.size(), 1));
  lr.cb.alloc(8 * std::max<size_t>(cb.size(), 1));
  lr.cstart.alloc(4 * cs.size());
  GRE_HIP(hipMemsetAsync(lr.acc.p, 0, 8 * std::max<size_t>(n, 1), s));
  if (n) GRE_HIP(hipMemcpyAsync(lr.rows.p, hr.data(), 4 * (size_t)n, hipMemcpyHostToDevice, s));
  if (!cv.empty()) {
    GRE_HIP(hipMemcpyAsync(lr.cv.p, cv.data(), 4 * cv.size(), hipMemcpyHostToDevice, s));
    GRE_HIP(hipMemcpyAsync(lr.cb.p, cb.data(), 8 * cb.size(), hipMemcpyHostToDevice, s));
  }
  GRE_HIP(hipMemcpyAsync(lr.cstart.p, cs.data(), 4 * cs.size(), hipMemcpyHostToDevice, s));
  GRE_HIP(hipStreamSynchronize(s));
  return lr;
}

// clears the frontier masks of the queued vertices
__global__ void k_ms_clear(const uint32_t* __restrict__ q, const MsCtl* __restrict__ c,
                           unsigned long long* __restrict__ fr) {
  if (c->mode == kMsDone) return;
  const uint32_t n = c->qn;
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x)
    fr[q[j]] = 0ULL;
}

__global__ void k_ms_edges(const unsigned long long* __restrict__ vis, const uint64_t* __restrict__ off, uint32_t nV,
                           unsigned long long* __restrict__ out) {
  uint64_t s = 0;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV; v += (uint64_t)gridDim.x * blockDim.x)
    s += (off[v + 1] - off[v]) * (uint64_t)__popcll(vis[v]);
  block_add(static_cast<unsigned long long>(s), out);
}

// Rows above kBcLong entries are cut into chunks of kBcChunk (one wave each):
// one wave walking a hub's tens of thousands of neighbours set the time of a
// whole level.  A chunk leaves per-lane partial sums; the row's finish adds
// them in chunk order (path counts are integers, exact in any order).
constexpr uint64_t kBcLong = 512;
constexpr uint64_t kBcChunk = 512;
// Those are the backward pass's; the forward pass's pair (whose chunking
// leaves the exact integer path counts unchanged) is picked per graph by
// timing these candidates on the first call.  The backward pass keeps its
// pair: its chunk partial sums fix the rounding of the dependencies.
constexpr uint64_t kBcCand[][2] = {{512, 512}, {256, 256}, {2048, 2048}, {1024, 4096}};

// forward fold of entries [b, e) of v's in-row: each lane adds the sigma of
// the in-neighbours that reach v for its source; gotl collects those lanes
__device__ inline void bc_fwd_run(const uint32_t* __restrict__ cidx, uint64_t b, uint64_t e,
                                  unsigned long long seen, const unsigned long long* __restrict__ fr,
                                  const double* __restrict__ sigma, int lane, double& acc,
                                  unsigned long long& gotl) {
  for (uint64_t base = b; base < e; base += 64) {
    // 64 in-neighbours and their frontier masks at once
    const uint64_t p = base + lane;
    const uint32_t w = p < e ? cidx[p] : 0u;
    const unsigned long long f = p < e ? (fr[w] & ~seen) : 0ULL;
    gotl |= f;
    unsigned long long act = __ballot(f != 0ULL);
    // sigma rows of the contributing neighbours, 8 gathers in flight, added
    // in row order
    while (act) {
      uint32_t wj[8];
      unsigned long long fj[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int j = act ? __ffsll(act) - 1 : 0;
        fj[t] = act ? __shfl(f, j, 64) : 0ULL;
        wj[t] = __shfl(w, j, 64);
        act &= act - 1ULL;
      }
      double sv[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) sv[t] = sigma[(uint64_t)wj[t] * 64 + lane];
#pragma unroll
      for (int t = 0; t < 8; ++t)
        if ((fj[t] >> lane) & 1ULL) acc += sv[t];
    }
  }
}

// backward fold of entries [b, e) of v's out-row: each lane adds q of the
// out-neighbours one level down for its source (a = v's lanes at this level)
__device__ inline void bc_bwd_run(const uint32_t* __restrict__ idx, uint64_t b, uint64_t e, unsigned long long a,
                                  const unsigned long long* __restrict__ nxt, const double* __restrict__ q, int lane,
                                  double& acc) {
  for (uint64_t base = b; base < e; base += 64) {
    // 64 out-neighbours and the lanes they hold at L+1, restricted to v's lanes at L
    const uint64_t p = base + lane;
    const uint32_t w = p < e ? idx[p] : 0u;
    const unsigned long long m = p < e ? (nxt[w] & a) : 0ULL;
    unsigned long long act = __ballot(m != 0ULL);
    while (act) {
      uint32_t wj[8];
      unsigned long long mj[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int j = act ? __ffsll(act) - 1 : 0;
        mj[t] = act ? __shfl(m, j, 64) : 0ULL;
        wj[t] = __shfl(w, j, 64);
        act &= act - 1ULL;
      }
      double qv[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) qv[t] = q[(uint64_t)wj[t] * 64 + lane];
#pragma unroll
      for (int t = 0; t < 8; ++t)
        if ((mj[t] >> lane) & 1ULL) acc += qv[t];
    }
  }
}

// The frontier of forward level L is q[L & 1] with lane masks fr[L & 1]; the
// level builds q / fr of the other parity.  Backward level l reads the lane
// masks of levels l and l + 1 from msk[l & 1] and msk[(l + 1) & 1].
template <typename T>
__device__ inline T* cur_of(const BcCtl* c, T* a0, T* a1) { return (c->L & 1) ? a1 : a0; }
template <typename T>
__device__ inline T* nxt_of(const BcCtl* c, T* a0, T* a1) { return (c->L & 1) ? a0 : a1; }

// forward level over the long rows' chunks: per chunk, each lane's partial
// sum and the lanes found
__global__ __launch_bounds__(kB) void k_bc_fwd_chunks(const uint64_t* __restrict__ coff,
                                                      const uint32_t* __restrict__ cidx,
                                                      const uint32_t* __restrict__ rows,
                                                      const uint32_t* __restrict__ cv,
                                                      const uint64_t* __restrict__ cb, uint32_t nchunks, uint64_t chunk,
                                                      const BcCtl* __restrict__ c, const unsigned long long* fr0,
                                                      const unsigned long long* fr1,
                                                      const unsigned long long* __restrict__ vis,
                                                      const double* __restrict__ sigma, double* __restrict__ psum,
                                                      unsigned long long* __restrict__ pgot) {
  if (!bu_on(c)) return;
  const unsigned long long* fr = cur_of(c, fr0, fr1);
  const int lane = threadIdx.x & 63;
  const unsigned long long full = c->full;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t k = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; k < nchunks; k += nw) {
    const uint32_t v = rows[cv[k]];
    const unsigned long long seen = vis[v];
    double acc = 0.0;
    unsigned long long gotl = 0;
    if (seen != full) {  // the same chunk, so the same answer, on every lane
      const uint64_t b = cb[k], rowend = coff[v + 1];
      bc_fwd_run(cidx, b, b + chunk < rowend ? b + chunk : rowend, seen, fr, sigma, lane, acc, gotl);
    }
    psum[k * 64 + lane] = acc;
    const unsigned long long got = wave_or(gotl);
    if (lane == 0) pgot[k] = got;
  }
}

// the long rows' forward results from their chunks (as k_bc_fwd does for the rest)
__global__ __launch_bounds__(kB) void k_bc_fwd_long_fin(const uint32_t* __restrict__ rows,
                                                        const uint32_t* __restrict__ cstart, uint32_t n,
                                                        const BcCtl* __restrict__ c,
                                                        const double* __restrict__ psum,
                                                        const unsigned long long* __restrict__ pgot,
                                                        unsigned long long* fr0, unsigned long long* fr1,
                                                        unsigned long long* __restrict__ vis,
                                                        double* __restrict__ sigma, uint8_t* __restrict__ lev) {
  if (!bu_on(c)) return;
  unsigned long long* nfr = nxt_of(c, fr0, fr1);
  const int lane = threadIdx.x & 63;
  const unsigned long long full = c->full;
  const uint8_t lnext = (uint8_t)(c->L + 1);
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; i < n; i += nw) {
    const uint32_t v = rows[i];
    const unsigned long long seen = vis[v];
    if (seen == full) { if (lane == 0) nfr[v] = 0; continue; }
    double acc = 0.0;
    unsigned long long got = 0;
    for (uint32_t k = cstart[i]; k < cstart[i + 1]; ++k) {
      acc += psum[(uint64_t)k * 64 + lane];
      got |= pgot[k];
    }
    if (lane == 0) nfr[v] = got;
    if (got) {
      if ((got >> lane) & 1ULL) {
        sigma[(uint64_t)v * 64 + lane] = acc;
        lev[(uint64_t)v * 64 + lane] = lnext;
      }
      if (lane == 0) vis[v] = seen | got;
    }
  }
}

// backward level over the long rows' chunks: each lane's partial sum
__global__ __launch_bounds__(kB) void k_bc_bwd_chunks(const uint64_t* __restrict__ off,
                                                      const uint32_t* __restrict__ idx,
                                                      const uint32_t* __restrict__ rows,
                                                      const uint32_t* __restrict__ cv,
                                                      const uint64_t* __restrict__ cb, uint32_t nchunks,
                                                      const BcCtl* __restrict__ c, const double* __restrict__ q,
                                                      const unsigned long long* m0, const unsigned long long* m1,
                                                      double* __restrict__ psum) {
  if (c->phase != kBcBwd) return;
  const unsigned long long* atl = cur_of(c, m0, m1);
  const unsigned long long* nxt = nxt_of(c, m0, m1);
  const int lane = threadIdx.x & 63;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t k = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; k < nchunks; k += nw) {
    const uint32_t v = rows[cv[k]];
    const unsigned long long a = atl[v];
    double acc = 0.0;
    if (a) {
      const uint64_t b = cb[k], rowend = off[v + 1];
      bc_bwd_run(idx, b, b + kBcChunk < rowend ? b + kBcChunk : rowend, a, nxt, q, lane, acc);
    }
    psum[k * 64 + lane] = acc;
  }
}

// the long rows' delta and q from their chunks' partial sums, in chunk order
__global__ __launch_bounds__(kB) void k_bc_bwd_long_fin(const uint32_t* __restrict__ rows,
                                                        const uint32_t* __restrict__ cstart, uint32_t n,
                                                        const BcCtl* __restrict__ c,
                                                        const double* __restrict__ psum,
                                                        const double* __restrict__ sigma,
                                                        double* __restrict__ delta, double* __restrict__ q,
                                                        const unsigned long long* m0, const unsigned long long* m1) {
  if (c->phase != kBcBwd) return;
  const unsigned long long* atl = cur_of(c, m0, m1);
  const int lane = threadIdx.x & 63;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t i = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; i < n; i += nw) {
    const uint32_t v = rows[i];
    const unsigned long long a = atl[v];
    if (!((a >> lane) & 1ULL)) continue;
    double acc = 0.0;
    for (uint32_t k = cstart[i]; k < cstart[i + 1]; ++k) acc += psum[(uint64_t)k * 64 + lane];
    const uint64_t o = (uint64_t)v * 64 + lane;
    const double sv = sigma[o];
    const double d = sv * acc;
    delta[o] = d;
    q[o] = (1.0 + d) / sv;
  }
}

// ---------------------------------------------------------------------------
// Brandes BC, 64 sources per chunk, one wave per vertex (lane = source)
// ---------------------------------------------------------------------------
// forward level L -> L+1: sigma_v[j] = sum over in-neighbours w at level L of sigma_w[j]
__global__ __launch_bounds__(kB) void k_bc_fwd(const uint64_t* __restrict__ coff, const uint32_t* __restrict__ cidx,
                                               uint32_t nV, uint64_t lng, const BcCtl* __restrict__ c,
                                               unsigned long long* fr0, unsigned long long* fr1,
                                               unsigned long long* __restrict__ vis, double* __restrict__ sigma,
                                               uint8_t* __restrict__ lev) {
  if (!bu_on(c)) return;
  const unsigned long long* fr = cur_of(c, fr0, fr1);
  unsigned long long* nfr = nxt_of(c, fr0, fr1);
  const int lane = threadIdx.x & 63;
  const unsigned long long full = c->full;
  const uint8_t lnext = (uint8_t)(c->L + 1);
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t v = wid; v < nV; v += nw) {
    const uint64_t b = coff[v], e = coff[v + 1];
    if (e - b > lng) continue;  // k_bc_fwd_chunks + k_bc_fwd_long_fin
    const unsigned long long seen = vis[v];
    if (seen == full) { if (lane == 0) nfr[v] = 0; continue; }
    double acc = 0.0;
    unsigned long long gotl = 0;
    bc_fwd_run(cidx, b, e, seen, fr, sigma, lane, acc, gotl);
    const unsigned long long got = wave_or(gotl);
    if (lane == 0) nfr[v] = got;
    if (got) {
      if ((got >> lane) & 1ULL) {
        sigma[(uint64_t)v * 64 + lane] = acc;
        lev[(uint64_t)v * 64 + lane] = lnext;
      }
      if (lane == 0) vis[v] = seen | got;
    }
  }
}

// backward level l: delta_v[j] = sigma_v[j] * sum over out-neighbours w at l+1 of q_w[j], where
// q_w[j] = (1+delta_w[j])/sigma_w[j] was stored when w's level was done; v's own q follows.
// atl[v] / nxt[v] = lanes of v at level l / l+1 (at the deepest level nxt is all zero:
// delta = 0 and q = 1/sigma)
__global__ __launch_bounds__(kB) void k_bc_bwd(const uint64_t* __restrict__ off, const uint32_t* __restrict__ idx,
                                               uint32_t nV, const BcCtl* __restrict__ c,
                                               const double* __restrict__ sigma, double* __restrict__ delta,
                                               double* __restrict__ q, const unsigned long long* m0,
                                               const unsigned long long* m1) {
  if (c->phase != kBcBwd) return;
  const unsigned long long* atl = cur_of(c, m0, m1);
  const unsigned long long* nxt = nxt_of(c, m0, m1);
  const int lane = threadIdx.x & 63;
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t v = wid; v < nV; v += nw) {
    const unsigned long long a = atl[v];
    if (!a) continue;
    const uint64_t b = off[v], e = off[v + 1];
    if (e - b > kBcLong) continue;  // k_bc_bwd_chunks + k_bc_bwd_long_fin
    const bool mine = (a >> lane) & 1ULL;
    double acc = 0.0;
    bc_bwd_run(idx, b, e, a, nxt, q, lane, acc);
    if (mine) {
      const uint64_t o = (uint64_t)v * 64 + lane;
      const double sv = sigma[o];
      const double d = sv * acc;
      delta[o] = d;
      q[o] = (1.0 + d) / sv;
    }
  }
}

// forward, top-down level L -> L+1 over the frontier's out-edges (edge-balanced
// like k_ms_td): v gains the lanes nw = fr[u] & ~vis[v], and sigma_v[j] +=
// sigma_u[j] for each of them -- path counts are integer-valued doubles below
// 2^53, so the atomic sums are exact in any order; the first edge that reaches
// v queues it.  vis stays the pre-level mask until k_bc_td_fin.
__global__ __launch_bounds__(kB) void k_bc_td(const uint64_t* __restrict__ off, const uint32_t* __restrict__ idx,
                                              uint32_t* q0, uint32_t* q1, const uint64_t* __restrict__ qpos,
                                              BcCtl* __restrict__ c, unsigned long long* fr0,
                                              unsigned long long* fr1, const unsigned long long* __restrict__ vis,
                                              double* __restrict__ sigma) {
  __shared__ uint32_t range[2];
  if (!td_on(c)) return;
  const uint32_t* q = cur_of(c, q0, q1);
  uint32_t* nq = nxt_of(c, q0, q1);
  const unsigned long long* fr = cur_of(c, fr0, fr1);
  unsigned long long* nfr = nxt_of(c, fr0, fr1);
  const uint32_t qn = c->qn;
  const uint64_t T = c->T;
  uint32_t* ncount = &c->nn;
  // edge blocks of kEB, grid-stride (the same trip count on every thread of a block)
  for (uint64_t blk = blockIdx.x; blk * kEB < T; blk += gridDim.x) {
    const uint64_t k0 = blk * kEB;
    const uint64_t k1 = k0 + kEB < T ? k0 + kEB : T;
    if (threadIdx.x < 2) {
      const uint64_t key = threadIdx.x == 0 ? k0 : k1 - 1;
      uint32_t lo = 0, hi = qn;
      while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (qpos[mid] <= key) lo = mid; else hi = mid;
      }
      range[threadIdx.x] = lo;
    }
    __syncthreads();
    const uint32_t i0 = range[0], i1 = range[1];
    for (int round = 0; round < kEB / kB; ++round) {
      const uint64_t k = k0 + (uint64_t)round * kB + threadIdx.x;
      bool take = false;
      uint32_t v = 0;
      if (k < k1) {
        uint32_t lo = i0, hi = i1 + 1;
        while (hi - lo > 1) {
          const uint32_t mid = lo + (hi - lo) / 2;
          if (qpos[mid] <= k) lo = mid; else hi = mid;
        }
        const uint32_t u = q[lo];
        v = idx[off[u] + (k - qpos[lo])];
        const unsigned long long nw = fr[u] & ~vis[v];
        if (nw) {
          take = atomicOr(nfr + v, nw) == 0ULL;
          unsigned long long m = nw;
          while (m) {
            const int j = __ffsll(m) - 1;
            atomicAdd(sigma + (uint64_t)v * 64 + j, sigma[(uint64_t)u * 64 + j]);
            m &= m - 1ULL;
          }
        }
      }
      block_append(take, v, nq, ncount);
    }
    __syncthreads();   // range[] is rewritten by the next edge block
  }
}

// after a top-down level: the queued vertices take their new lanes at level L+1;
// sums[0] += their out-degrees
__global__ void k_bc_td_fin(BcCtl* __restrict__ c, const uint32_t* q0, const uint32_t* q1,
                            const unsigned long long* fr0, const unsigned long long* fr1,
                            unsigned long long* __restrict__ vis, uint8_t* __restrict__ lev,
                            const uint64_t* __restrict__ off) {
  if (!td_on(c)) return;
  const uint32_t* nq = nxt_of(c, q0, q1);
  const unsigned long long* nfr = nxt_of(c, fr0, fr1);
  const uint32_t nn = c->nn;
  const uint8_t lnext = (uint8_t)(c->L + 1);
  uint64_t deg = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = nq[i];
    unsigned long long m = nfr[v];
    vis[v] |= m;
    while (m) {
      const int j = __ffsll(m) - 1;
      lev[(uint64_t)v * 64 + j] = lnext;
      m &= m - 1ULL;
    }
    deg += off[v + 1] - off[v];
  }
  block_add(static_cast<unsigned long long>(deg), c->sums);
}

// after a bottom-up level: the new frontier (nonzero masks) as a queue, with
// sums[0] += its out-degrees
__global__ __launch_bounds__(kB) void k_bc_queue(uint32_t nV, BcCtl* __restrict__ c, const unsigned long long* fr0,
                                                 const unsigned long long* fr1, const uint64_t* __restrict__ off,
                                                 uint32_t* q0, uint32_t* q1) {
  if (!bu_on(c)) return;
  const unsigned long long* nfr = nxt_of(c, fr0, fr1);
  uint32_t* nq = nxt_of(c, q0, q1);
  uint64_t deg = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < nV; base += stride) {
    const uint64_t v = base + threadIdx.x;
    const bool take = v < nV && nfr[v] != 0ULL;
    if (take) deg += off[v + 1] - off[v];
    block_append(take, (uint32_t)v, nq, &c->nn);
  }
  block_add(static_cast<unsigned long long>(deg), c->sums);
}

// top-down levels: out-degrees of the frontier (then their prefix by the
// device-length scan)
__global__ void k_bc_qdeg(const BcCtl* __restrict__ c, const uint32_t* q0, const uint32_t* q1,
                          const uint64_t* __restrict__ off, uint64_t* __restrict__ d) {
  if (!td_on(c)) return;
  const uint32_t* q = cur_of(c, q0, q1);
  const uint32_t n = c->qn;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i <= n; i += (uint64_t)gridDim.x * blockDim.x)
    d[i] = i < n ? off[q[i] + 1] - off[q[i]] : 0;
}

// end of a forward level: the frontier's masks are cleared (fr of either
// parity is zero outside its own queue, which a top-down level relies on)
__global__ void k_bc_clear(const BcCtl* __restrict__ c, const uint32_t* q0, const uint32_t* q1,
                           unsigned long long* fr0, unsigned long long* fr1) {
  if (c->phase != kBcFwd) return;
  const uint32_t* q = cur_of(c, q0, q1);
  unsigned long long* fr = cur_of(c, fr0, fr1);
  const uint32_t n = c->qn;
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x)
    fr[q[j]] = 0ULL;
}

// backward level l: mask of the lanes whose level is l into msk[l & 1]; at the
// deepest level msk[(l + 1) & 1] (level l + 1, empty) is cleared too
__global__ void k_bc_mask(const BcCtl* __restrict__ c, uint32_t nV, const uint8_t* __restrict__ lev,
                          unsigned long long* m0, unsigned long long* m1) {
  if (c->phase != kBcBwd) return;
  unsigned long long* m = cur_of(c, m0, m1);
  unsigned long long* other = nxt_of(c, m0, m1);
  const uint8_t l = (uint8_t)c->L;
  const bool deepest = c->L == c->depth;
  const int lane = threadIdx.x & 63;
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t v = wid; v < nV; v += nw) {
    const unsigned long long b = __ballot(lev[v * 64 + lane] == l);
    if (lane == 0) {
      m[v] = b;
      if (deepest) other[v] = 0ULL;
    }
  }
}

// after backward level 0: bc[v] += sum over the chunk's lanes of delta (the
// lane's own source excluded)
__global__ void k_bc_accum(const BcCtl* __restrict__ c, uint32_t nV, const uint32_t* __restrict__ src,
                           const double* __restrict__ delta, double* __restrict__ bc) {
  if (c->phase != kBcBwd || c->L != 0) return;
  const int nc = (int)c->nc;
  const uint32_t* cs = src + (uint64_t)c->chunk * 64;
  const int lane = threadIdx.x & 63;
  const uint64_t wid = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t v = wid; v < nV; v += nw) {
    double d = (lane < nc && cs[lane] != v) ? delta[v * 64 + lane] : 0.0;
    d = wave_sum(d);
    if (lane == 0) bc[v] += d;
  }
}

// set-up of a chunk: its per-source arrays cleared
__global__ void k_bc_init(const BcCtl* __restrict__ c, uint32_t nV, double* __restrict__ sigma,
                          double* __restrict__ delta, uint8_t* __restrict__ lev, unsigned long long* __restrict__ vis,
                          unsigned long long* __restrict__ fr0, unsigned long long* __restrict__ fr1) {
  if (c->phase != kBcInit) return;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < (uint64_t)nV * 64;
       i += (uint64_t)gridDim.x * blockDim.x) {
    sigma[i] = 0.0;
    delta[i] = 0.0;
    lev[i] = 0xff;
    if ((i & 63) == 0) { vis[i >> 6] = 0; fr0[i >> 6] = 0; fr1[i >> 6] = 0; }
  }
}

// first kernel of a step: a chunk's seeds and first frontier (its distinct
// sources, uq / uqn, prepared by the host), then the forward level's direction
__global__ void k_bc_begin(BcCtl* __restrict__ c, uint64_t nE, const uint64_t* __restrict__ off,
                           const uint32_t* __restrict__ src, const uint32_t* __restrict__ uq,
                           const uint32_t* __restrict__ uqn, double* __restrict__ sigma, uint8_t* __restrict__ lev,
                           unsigned long long* __restrict__ vis, unsigned long long* __restrict__ fr0,
                           uint32_t* __restrict__ q0) {
  if (c->phase == kBcInit) {
    const uint32_t c0 = c->chunk * 64;
    const uint32_t nc = c->nsrc - c0 < 64 ? c->nsrc - c0 : 64;
    for (uint32_t j = 0; j < nc; ++j) {
      const uint64_t s = src[c0 + j];
      sigma[s * 64 + j] = 1.0;
      lev[s * 64 + j] = 0;
      vis[s] |= 1ULL << j;
      fr0[s] |= 1ULL << j;
    }
    const uint32_t nu = uqn[c->chunk];
    uint64_t mf = 0;
    for (uint32_t i = 0; i < nu; ++i) {
      const uint32_t v = uq[c0 + i];
      q0[i] = v;
      mf += off[v + 1] - off[v];
    }
    c->nc = nc;
    c->full = nc == 64 ? ~0ULL : ((1ULL << nc) - 1);
    c->qn = nu;
    c->mf = mf;
    c->L = 0;
    c->phase = kBcFwd;
  }
  if (c->phase != kBcFwd) return;
  if (c->L >= kBcMaxDepth) { c->phase = kBcTooDeep; return; }
  // top-down over the frontier's out-edges while they are few (under E / 14),
  // else the bottom-up pull over every vertex
  c->mode = c->mf * 14 < nE ? kMsTopDown : kMsBottomUp;
  c->nn = 0;
  c->T = 0;
  c->sums[0] = 0;
  c->sums[1] = 0;
}

// last kernel of a step: the next forward level, the turn to the backward
// pass once a level finds nothing, the next backward level, the next chunk
__global__ void k_bc_end(BcCtl* c) {
  if (c->phase == kBcFwd) {
    if (c->nn == 0) {
      c->phase = kBcBwd;
      c->depth = c->L;
    } else {
      c->qn = c->nn;
      c->mf = c->sums[0];
      c->L += 1;
    }
  } else if (c->phase == kBcBwd) {
    if (c->L == 0) {
      c->chunk += 1;
      c->phase = c->chunk < c->nchunks ? kBcInit : kBcDone;
    } else {
      c->L -= 1;
    }
  }
}

// after backward level 0: the chunk's lanes, lane-major (v*64 + j), into
// rows 64 chunk + j of the source-major output (j*nV + v)
template <typename T, typename U>
__global__ void k_bc_export(const BcCtl* __restrict__ c, uint32_t nV, const T* __restrict__ in, U* __restrict__ out,
                            bool lev8) {
  if (c->phase != kBcBwd || c->L != 0) return;
  const uint64_t nc = c->nc;
  U* o = out + (uint64_t)c->chunk * 64 * nV;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < (uint64_t)nV * nc;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t j = i / nV, v = i % nV;
    const T x = in[v * 64 + j];
    if (lev8) o[i] = (U)(x == (T)0xff ? -1 : (int)x);
    else o[i] = (U)x;
  }
}

// Working arrays of the batched BC, grown to the largest graph, source count
// and long-row chunk count seen and kept for the process (like the BFS ones):
// a call allocates only the per-source outputs it is asked for.
struct BcBufs {
  DevBuf sigma, delta, qv, lev, vis, fr[2], msk[2], q[2], qdeg, qpos, bsum, bc, ctl, psum, pgot, src, uq, uqn;
  uint32_t nV = 0, nchunks = 0, pchunks = 0;
  BcCtl* ctl_host = nullptr;   // pinned
};
BcBufs& bc_bufs(uint32_t nV, uint32_t nchunks, uint32_t pchunks) {
  static BcBufs& b = *new BcBufs();  // lives until the process ends
  if (b.nV < nV || !b.sigma.p) {
    const size_t n64 = (size_t)nV * 64;
    b.sigma.alloc(8 * n64);
    b.delta.alloc(8 * n64);
    b.qv.alloc(8 * n64);
    b.lev.alloc(n64);
    b.vis.alloc(8 * (size_t)nV);
    for (int k = 0; k < 2; ++k) {
      b.fr[k].alloc(8 * (size_t)nV);
      b.msk[k].alloc(8 * (size_t)nV);
      b.q[k].alloc(4 * (size_t)nV + 4 * 64);
    }
    b.qdeg.alloc(8 * ((size_t)nV + 1));
    b.qpos.alloc(8 * ((size_t)nV + 1));
    b.bsum.alloc(8 * (((size_t)nV + 1 + kScanB - 1) / kScanB + 1));
    b.bc.alloc(8 * (size_t)nV);
    b.nV = nV;
  }
  if (b.nchunks < nchunks) {
    b.src.alloc(4 * 64 * (size_t)nchunks);
    b.uq.alloc(4 * 64 * (size_t)nchunks);
    b.uqn.alloc(4 * (size_t)nchunks);
    b.nchunks = nchunks;
  }
  if (b.pchunks < pchunks) {
    b.psum.alloc(8 * 64 * (size_t)pchunks);
    b.pgot.alloc(8 * (size_t)pchunks);
    b.pchunks = pchunks;
  }
  if (!b.ctl.p) b.ctl.alloc(sizeof(BcCtl));
  if (!b.ctl_host) GRE_HIP(hipHostMalloc((void**)&b.ctl_host, sizeof(BcCtl), hipHostMallocDefault));
  return b;
}

// One 64-source BFS over g from srcs (uq: their distinct ids, the first
// frontier), the bottom-up pass cutting rows above lng in-edges into chunks of
// `chunk`; levels into lv when non-null.  Leaves the visited masks in the
// working arrays.
void ms_run(gre_graph* g, const uint32_t* srcs, int nsrc, const std::vector<uint32_t>& uq, int32_t* lv,
            uint64_t lng, uint64_t chunk) {
  hipStream_t s = stream();
  const uint32_t nV = g->nV;
  const uint64_t* off = g->off.as<uint64_t>();
  MsBufs& mb = ms_bufs(nV);
  DevBuf* fr = mb.fr;
  DevBuf* q = mb.q;
  DevBuf &vis = mb.vis, &nxt = mb.nxt, &qdeg = mb.qdeg, &qpos = mb.qpos;
  LongRows& lr = long_rows(g, true, lng, chunk, s);
  GRE_HIP(hipMemsetAsync(vis.p, 0, 8 * (size_t)nV, s));
  GRE_HIP(hipMemsetAsync(lr.acc.p, 0, 8 * std::max<size_t>(lr.n, 1), s));
  GRE_HIP(hipMemsetAsync(fr[0].p, 0, 8 * (size_t)nV, s));
  GRE_HIP(hipMemsetAsync(fr[1].p, 0, 8 * (size_t)nV, s));
  GRE_HIP(hipMemsetAsync(nxt.p, 0, 8 * (size_t)nV, s));
  GRE_HIP(hipMemcpyAsync(mb.src.p, srcs, 4 * (size_t)nsrc, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_ms_seed, dim3(1), dim3(1), 0, s, mb.src.as<uint32_t>(), nsrc, (unsigned long long*)vis.p,
                     (unsigned long long*)fr[0].p, lv, nV);
  GRE_HIP(hipGetLastError());
  GRE_HIP(hipMemcpyAsync(q[0].p, uq.data(), 4 * uq.size(), hipMemcpyHostToDevice, s));
  GRE_HIP(hipStreamSynchronize(s));
  const unsigned long long full = nsrc == 64 ? ~0ULL : ((1ULL << nsrc) - 1);
  // level 0: the distinct sources, top-down, mu = every edge
  MsCtl* hc = mb.ctl_host;
  *hc = MsCtl{};
  hc->qn = (uint32_t)uq.size();
  hc->mode = kMsTopDown;
  hc->mu = g->nE;
  MsCtl* dc = mb.ctl.as<MsCtl>();
  GRE_HIP(hipMemcpyAsync(dc, hc, sizeof(MsCtl), hipMemcpyHostToDevice, s));
  const uint64_t nbmax = ((uint64_t)nV + 1 + kScanB - 1) / kScanB;
  GRE_CHECK(nbmax <= 0x7fffffffULL, GRE_DEV_FAILURE, "graph too large for the multi-source BFS");
  const unsigned tdgrid = (unsigned)std::min<uint64_t>(std::max<uint64_t>(g->nE / kEB, 1), 2 * kMsGrid);
  const unsigned vgrid = std::min<unsigned>(gcap((uint64_t)nV + 1), kMsGrid);
  const unsigned sgrid = (unsigned)std::min<uint64_t>(nbmax, kMsScanGrid);
  // one level: control, then both directions' kernels (the one not chosen
  // returns at once), the old frontier's masks cleared, control again
  auto level = [&](int cur) {
    const int nx = cur ^ 1;
    hipLaunchKernelGGL(k_ms_ctl_begin, dim3(1), dim3(1), 0, s, dc, nV);
    hipLaunchKernelGGL(k_ms_qdeg, dim3(vgrid), dim3(kB), 0, s, dc, q[cur].as<uint32_t>(), off, qdeg.as<uint64_t>());
    hipLaunchKernelGGL(k_ms_scan_sums<MsCtl>, dim3(sgrid), dim3(kB), 0, s, dc, qdeg.as<uint64_t>(),
                       mb.bsum.as<uint64_t>());
    hipLaunchKernelGGL(k_ms_scan_top<MsCtl>, dim3(1), dim3(1024), 0, s, dc, mb.bsum.as<uint64_t>());
    hipLaunchKernelGGL(k_ms_scan_apply<MsCtl>, dim3(sgrid), dim3(kB), 0, s, dc, qdeg.as<uint64_t>(),
                       mb.bsum.as<uint64_t>(), qpos.as<uint64_t>());
    hipLaunchKernelGGL(k_ms_td, dim3(tdgrid), dim3(kB), 0, s, off, g->idx.as<uint32_t>(), q[cur].as<uint32_t>(),
                       qpos.as<uint64_t>(), dc, (const unsigned long long*)fr[cur].p, (const unsigned long long*)vis.p,
                       (unsigned long long*)nxt.p, q[nx].as<uint32_t>());
    hipLaunchKernelGGL(k_ms_td_fin, dim3(vgrid), dim3(kB), 0, s, q[nx].as<uint32_t>(), dc, full,
                       (unsigned long long*)nxt.p, (unsigned long long*)fr[nx].p, (unsigned long long*)vis.p, off, lv,
                       nV);
    // the long rows' chunks read vis and fr[cur] only; k_ms_bu skips those rows
    if (lr.nchunks)
      hipLaunchKernelGGL(k_ms_bu_chunks, dim3(std::min<uint64_t>(((uint64_t)lr.nchunks + 3) / 4, kMsGrid)), dim3(kB),
                         0, s, g->coff.as<uint64_t>(), g->cidx.as<uint32_t>(), lr.rows.as<uint32_t>(),
                         lr.cv.as<uint32_t>(), lr.cb.as<uint64_t>(), lr.nchunks, full,
                         (const unsigned long long*)fr[cur].p, (const unsigned long long*)vis.p,
                         (unsigned long long*)lr.acc.p, dc, chunk);
    hipLaunchKernelGGL(k_ms_bu, dim3(vgrid), dim3(kB), 0, s, g->coff.as<uint64_t>(), g->cidx.as<uint32_t>(), off, nV,
                       full, (const unsigned long long*)fr[cur].p, (unsigned long long*)fr[nx].p,
                       (unsigned long long*)vis.p, lv, dc, q[nx].as<uint32_t>(), lng);
    if (lr.n)
      hipLaunchKernelGGL(k_ms_bu_long_fin, dim3(std::min<unsigned>(gcap(lr.n), kMsGrid)), dim3(kB), 0, s,
                         lr.rows.as<uint32_t>(), lr.n, full, (unsigned long long*)lr.acc.p,
                         (unsigned long long*)fr[nx].p, (unsigned long long*)vis.p, off, lv, nV, dc,
                         q[nx].as<uint32_t>());
    hipLaunchKernelGGL(k_ms_clear, dim3(vgrid), dim3(kB), 0, s, q[cur].as<uint32_t>(), dc,
                       (unsigned long long*)fr[cur].p);
    hipLaunchKernelGGL(k_ms_ctl_end, dim3(1), dim3(1), 0, s, dc);
  };
  // kMsBatch levels captured once per launch configuration and replayed
  // until the control says done: one read-back per batch, none per level
  static std::map<std::vector<uint64_t>, hipGraphExec_t> graphs;
  auto P = [](const void* p) { return (uint64_t)(uintptr_t)p; };
  const std::vector<uint64_t> key = {P(g),         nV,          g->nE,        full,         P(lv),        P(dc),
                                     P(vis.p),     P(fr[0].p),  P(fr[1].p),   P(nxt.p),     P(q[0].p),    P(q[1].p),
                                     P(qdeg.p),    P(qpos.p),   P(mb.bsum.p), P(lr.acc.p),  P(lr.rows.p), P(lr.cv.p),
                                     P(lr.cb.p),   lr.n,        lr.nchunks,   lng,          chunk,        P(off),
                                     P(g->idx.p),  P(g->coff.p), P(g->cidx.p)};
  hipGraphExec_t exec = nullptr;
  auto it = graphs.find(key);
  if (it != graphs.end()) {
    exec = it->second;
  } else {
    // calls that return levels use a new output buffer each time: keep the
    // cache small (the stream is idle here: no replay is in flight)
    if (graphs.size() >= 16) {
      GRE_HIP(hipStreamSynchronize(s));
      for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
      graphs.clear();
    }
    hipGraph_t graph = nullptr;
    GRE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < kMsBatch; ++k) level(k & 1);
    GRE_HIP(hipStreamEndCapture(s, &graph));
    GRE_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    graphs[key] = exec;
  }
  ProfScope ps("msbfs_levels");
  for (uint64_t done_levels = 0;; done_levels += kMsBatch) {
    GRE_CHECK(done_levels <= (uint64_t)nV + kMsBatch, GRE_DEV_FAILURE, "multi-source BFS did not terminate");
    GRE_HIP(hipGraphLaunch(exec, s));
    GRE_HIP(hipMemcpyAsync(&hc->mode, &dc->mode, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
    if (hc->mode == kMsDone) break;
  }
}

// Times fn() once on the library stream (after one untimed call that builds
// and captures what it needs), in ms.
template <typename F>
float time_twice(F&& fn) {
  hipStream_t s = stream();
  fn();
  hipEvent_t a, b;
  GRE_HIP(hipEventCreate(&a));
  GRE_HIP(hipEventCreate(&b));
  float ms = 0.f;
  try {
    GRE_HIP(hipEventRecord(a, s));
    fn();
    GRE_HIP(hipEventRecord(b, s));
    GRE_HIP(hipEventSynchronize(b));
    GRE_HIP(hipEventElapsedTime(&ms, a, b));
  } catch (...) {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    throw;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms;
}

// The first search of a graph picks the bottom-up pass's (threshold, chunk)
// pair: each candidate runs on this call's sources and the fastest is kept.
void ms_tune(gre_graph* g, const uint32_t* srcs, int nsrc, const std::vector<uint32_t>& uq) {
  float best = 1e30f;
  for (const auto& c : kMsCand) {
    const float ms = time_twice([&] { ms_run(g, srcs, nsrc, uq, nullptr, c[0], c[1]); });
    if (ms < best) {
      best = ms;
      g->ms_long = c[0];
      g->ms_chunk = c[1];
    }
  }
}

// Batched Brandes BC of every 64-source chunk of srcs in one device-driven
// launch sequence, the forward pass cutting rows above flong in-edges into
// chunks of fchunk; bc sums are left in the working array, per-source arrays
// go to xs / xd / xl (device, nsrc x nV) when non-null.
void bc_run(gre_graph* g, const uint32_t* srcs, int nsrc, double* xs, double* xd, int32_t* xl, uint64_t flong,
          uint64_t fchunk) {
  hipStream_t s = stream();
  const uint32_t nV = g->nV;
  const uint32_t nchunks = (uint32_t)((nsrc + 63) / 64);
  LongRows& lf = long_rows(g, true, flong, fchunk, s);
  LongRows& lb = long_rows(g, false, kBcLong, kBcChunk, s);
  BcBufs& b = bc_bufs(nV, nchunks, std::max<uint32_t>(std::max(lf.nchunks, lb.nchunks), 1));
  // each chunk's distinct sources: its first frontier
  std::vector<uint32_t> uq(64 * (size_t)nchunks, 0), uqn(nchunks, 0);
  for (uint32_t k = 0; k < nchunks; ++k) {
    const int c0 = 64 * (int)k, nc = std::min(64, nsrc - c0);
    std::vector<uint32_t> u(srcs + c0, srcs + c0 + nc);
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    std::copy(u.begin(), u.end(), uq.begin() + c0);
    uqn[k] = (uint32_t)u.size();
  }
  GRE_HIP(hipMemsetAsync(b.bc.p, 0, 8 * (size_t)nV, s));
  GRE_HIP(hipMemcpyAsync(b.src.p, srcs, 4 * (size_t)nsrc, hipMemcpyHostToDevice, s));
  GRE_HIP(hipMemcpyAsync(b.uq.p, uq.data(), 4 * uq.size(), hipMemcpyHostToDevice, s));
  GRE_HIP(hipMemcpyAsync(b.uqn.p, uqn.data(), 4 * uqn.size(), hipMemcpyHostToDevice, s));
  BcCtl* hc = b.ctl_host;
  *hc = BcCtl{};
  hc->phase = kBcInit;
  hc->nchunks = nchunks;
  hc->nsrc = (uint32_t)nsrc;
  BcCtl* dc = b.ctl.as<BcCtl>();
  GRE_HIP(hipMemcpyAsync(dc, hc, sizeof(BcCtl), hipMemcpyHostToDevice, s));

  const uint64_t nbmax = ((uint64_t)nV + 1 + kScanB - 1) / kScanB;
  GRE_CHECK(nbmax <= 0x7fffffffULL, GRE_DEV_FAILURE, "graph too large for gre_bc");
  const unsigned tdgrid = (unsigned)std::min<uint64_t>(std::max<uint64_t>(g->nE / kEB, 1), 2 * kMsGrid);
  const unsigned vgrid = std::min<unsigned>(gcap((uint64_t)nV + 1), kMsGrid);
  const unsigned sgrid = (unsigned)std::min<uint64_t>(nbmax, kMsScanGrid);
  auto wgrid = [](uint64_t waves) {
    return (unsigned)std::min<uint64_t>(std::max<uint64_t>((waves + 3) / 4, 1), kMsGrid);
  };
  const unsigned gw = wgrid(nV);
  const uint64_t* off = g->off.as<uint64_t>();
  const uint64_t* coff = g->coff.as<uint64_t>();
  const uint32_t* idx = g->idx.as<uint32_t>();
  const uint32_t* cidx = g->cidx.as<uint32_t>();
  double* sigma = b.sigma.as<double>();
  double* delta = b.delta.as<double>();
  double* qv = b.qv.as<double>();
  uint8_t* lev = b.lev.as<uint8_t>();
  auto* vis = (unsigned long long*)b.vis.p;
  auto* fr0 = (unsigned long long*)b.fr[0].p;
  auto* fr1 = (unsigned long long*)b.fr[1].p;
  auto* m0 = (unsigned long long*)b.msk[0].p;
  auto* m1 = (unsigned long long*)b.msk[1].p;
  uint32_t* q0 = b.q[0].as<uint32_t>();
  uint32_t* q1 = b.q[1].as<uint32_t>();
  // one step: set-up (chunk start), a forward level in either direction, a
  // backward level, the chunk's accumulation and outputs; all but the
  // step's own kernels return at once
  auto step = [&]() {
    hipLaunchKernelGGL(k_bc_init, dim3(std::min<unsigned>(gcap((uint64_t)nV * 64), kMsGrid)), dim3(kB), 0, s, dc,
                       nV, sigma, delta, lev, vis, fr0, fr1);
    hipLaunchKernelGGL(k_bc_begin, dim3(1), dim3(1), 0, s, dc, g->nE, off, b.src.as<uint32_t>(),
                       b.uq.as<uint32_t>(), b.uqn.as<uint32_t>(), sigma, lev, vis, fr0, q0);
    // forward, top-down
    hipLaunchKernelGGL(k_bc_qdeg, dim3(vgrid), dim3(kB), 0, s, dc, q0, q1, off, b.qdeg.as<uint64_t>());
    hipLaunchKernelGGL(k_ms_scan_sums<BcCtl>, dim3(sgrid), dim3(kB), 0, s, dc, b.qdeg.as<uint64_t>(),
                       b.bsum.as<uint64_t>());
    hipLaunchKernelGGL(k_ms_scan_top<BcCtl>, dim3(1), dim3(1024), 0, s, dc, b.bsum.as<uint64_t>());
    hipLaunchKernelGGL(k_ms_scan_apply<BcCtl>, dim3(sgrid), dim3(kB), 0, s, dc, b.qdeg.as<uint64_t>(),
                       b.bsum.as<uint64_t>(), b.qpos.as<uint64_t>());
    hipLaunchKernelGGL(k_bc_td, dim3(tdgrid), dim3(kB), 0, s, off, idx, q0, q1, b.qpos.as<uint64_t>(), dc, fr0, fr1,
                       vis, sigma);
    hipLaunchKernelGGL(k_bc_td_fin, dim3(vgrid), dim3(kB), 0, s, dc, q0, q1, fr0, fr1, vis, lev, off);
    // forward, bottom-up (the long rows' chunks read fr and vis only; k_bc_fwd skips those rows)
    if (lf.nchunks)
      hipLaunchKernelGGL(k_bc_fwd_chunks, dim3(wgrid(lf.nchunks)), dim3(kB), 0, s, coff, cidx,
                         lf.rows.as<uint32_t>(), lf.cv.as<uint32_t>(), lf.cb.as<uint64_t>(), lf.nchunks, fchunk, dc, fr0,
                         fr1, vis, sigma, b.psum.as<double>(), (unsigned long long*)b.pgot.p);
    hipLaunchKernelGGL(k_bc_fwd, dim3(gw), dim3(kB), 0, s, coff, cidx, nV, flong, dc, fr0, fr1, vis, sigma, lev);
    if (lf.n)
      hipLaunchKernelGGL(k_bc_fwd_long_fin, dim3(wgrid(lf.n)), dim3(kB), 0, s, lf.rows.as<uint32_t>(),
                         lf.cstart.as<uint32_t>(), lf.n, dc, b.psum.as<double>(),
                         (const unsigned long long*)b.pgot.p, fr0, fr1, vis, sigma, lev);
    hipLaunchKernelGGL(k_bc_queue, dim3(vgrid), dim3(kB), 0, s, nV, dc, fr0, fr1, off, q0, q1);
    hipLaunchKernelGGL(k_bc_clear, dim3(vgrid), dim3(kB), 0, s, dc, q0, q1, fr0, fr1);
    // backward
    hipLaunchKernelGGL(k_bc_mask, dim3(gw), dim3(kB), 0, s, dc, nV, lev, m0, m1);
    if (lb.nchunks)
      hipLaunchKernelGGL(k_bc_bwd_chunks, dim3(wgrid(lb.nchunks)), dim3(kB), 0, s, off, idx,
                         lb.rows.as<uint32_t>(), lb.cv.as<uint32_t>(), lb.cb.as<uint64_t>(), lb.nchunks, dc, qv, m0,
                         m1, b.psum.as<double>());
    hipLaunchKernelGGL(k_bc_bwd, dim3(gw), dim3(kB), 0, s, off, idx, nV, dc, sigma, delta, qv, m0, m1);
    if (lb.n)
      hipLaunchKernelGGL(k_bc_bwd_long_fin, dim3(wgrid(lb.n)), dim3(kB), 0, s, lb.rows.as<uint32_t>(),
                         lb.cstart.as<uint32_t>(), lb.n, dc, b.psum.as<double>(), sigma, delta, qv, m0, m1);
    hipLaunchKernelGGL(k_bc_accum, dim3(gw), dim3(kB), 0, s, dc, nV, b.src.as<uint32_t>(), delta,
                       b.bc.as<double>());
    const unsigned xg = std::min<unsigned>(gcap((uint64_t)nV * 64), kMsGrid);
    if (xs) hipLaunchKernelGGL((k_bc_export<double, double>), dim3(xg), dim3(kB), 0, s, dc, nV, sigma, xs, false);
    if (xd) hipLaunchKernelGGL((k_bc_export<double, double>), dim3(xg), dim3(kB), 0, s, dc, nV, delta, xd, false);
    if (xl) hipLaunchKernelGGL((k_bc_export<uint8_t, int32_t>), dim3(xg), dim3(kB), 0, s, dc, nV, lev, xl, true);
    hipLaunchKernelGGL(k_bc_end, dim3(1), dim3(1), 0, s, dc);
  };
  // kBcBatch steps captured once per launch configuration and replayed
  // until the control says done: one read-back per batch, none per level
  // or chunk
  static std::map<std::vector<uint64_t>, hipGraphExec_t> graphs;
  auto P = [](const void* p) { return (uint64_t)(uintptr_t)p; };
  const std::vector<uint64_t> key = {
      P(g),         nV,           g->nE,        P(dc),        P(sigma),      P(delta),     P(qv),
      P(lev),       P(vis),       P(fr0),       P(fr1),       P(m0),         P(m1),        P(q0),
      P(q1),        P(b.qdeg.p),  P(b.qpos.p),  P(b.bsum.p),  P(b.psum.p),   P(b.pgot.p),  P(b.bc.p),
      P(b.src.p),   P(b.uq.p),    P(b.uqn.p),   P(lf.rows.p), P(lf.cv.p),    P(lf.cb.p),   P(lf.cstart.p),
      lf.n,         lf.nchunks,   P(lb.rows.p), P(lb.cv.p),   P(lb.cb.p),    P(lb.cstart.p), lb.n,
      lb.nchunks,   P(off),       P(idx),       P(coff),      P(cidx),       P(xs),        P(xd),
      P(xl),        flong,        fchunk};
  hipGraphExec_t exec = nullptr;
  auto it = graphs.find(key);
  if (it != graphs.end()) {
    exec = it->second;
  } else {
    // calls that return per-source arrays use new output buffers each time:
    // keep the cache small (the stream is idle here: no replay is in flight)
    if (graphs.size() >= 16) {
      GRE_HIP(hipStreamSynchronize(s));
      for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
      graphs.clear();
    }
    hipGraph_t graph = nullptr;
    GRE_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < kBcBatch; ++k) step();
    GRE_HIP(hipStreamEndCapture(s, &graph));
    GRE_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    graphs[key] = exec;
  }
  {
    ProfScope ps("bc_steps");
    // a chunk takes at most 2 (kBcMaxDepth + 1) steps
    const uint64_t max_steps = (uint64_t)nchunks * 2 * (kBcMaxDepth + 1) + kBcBatch;
    for (uint64_t done_steps = 0;; done_steps += kBcBatch) {
      GRE_CHECK(done_steps <= max_steps, GRE_DEV_FAILURE, "gre_bc did not terminate");
      GRE_HIP(hipGraphLaunch(exec, s));
      GRE_HIP(hipMemcpyAsync(&hc->phase, &dc->phase, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      GRE_HIP(hipStreamSynchronize(s));
      if (hc->phase >= kBcDone) break;
    }
  }
  GRE_CHECK(hc->phase == kBcDone, GRE_DEV_FAILURE, "BFS depth above 250 levels is not supported by gre_bc");
}

// The first BC of a graph picks the forward pass's (threshold, chunk) pair:
// each candidate runs on this call's first chunk of sources and the fastest is
// kept.  Path counts are integers, exact in any order, so the pick changes no
// result.
void bc_tune(gre_graph* g, const uint32_t* srcs, int nsrc) {
  float best = 1e30f;
  for (const auto& c : kBcCand) {
    const float ms = time_twice([&] { bc_run(g, srcs, std::min(nsrc, 64), nullptr, nullptr, nullptr, c[0], c[1]); });
    if (ms < best) {
      best = ms;
      g->bc_flong = c[0];
      g->bc_fchunk = c[1];
    }
  }
}

}  // namespace
}  // namespace gre

using namespace gre;

extern "C" {

int gre_bfs_dev(gre_graph* g, uint32_t src, int32_t* level_dev, uint64_t* edges_out, int32_t* depth_out) {
  return guard([&] {
    GRE_CHECK(g, GRE_DEV_NULL, "null graph");
    GRE_CHECK(src < g->nV, GRE_DEV_NULL, "source id >= nV");
    int32_t* lev = level_dev ? level_dev : bfs_bufs(g->nV).lev.as<int32_t>();
    bfs_run(g, src, lev, edges_out, depth_out);
  });
}

int gre_bfs(gre_graph* g, uint32_t src, int32_t* level_out, uint64_t* edges_out, int32_t* depth_out) {
  return guard([&] {
    GRE_CHECK(g, GRE_DEV_NULL, "null graph");
    GRE_CHECK(src < g->nV, GRE_DEV_NULL, "source id >= nV");
    DevBuf& lev = bfs_bufs(g->nV).lev;
    bfs_run(g, src, lev.as<int32_t>(), edges_out, depth_out);
    if (level_out) {
      GRE_HIP(hipMemcpyAsync(level_out, lev.p, 4 * (size_t)g->nV, hipMemcpyDeviceToHost, stream()));
      GRE_HIP(hipStreamSynchronize(stream()));
    }
  });
}

int gre_msbfs(gre_graph* g, const uint32_t* srcs, int nsrc, int32_t* level_out, uint64_t* edges_out) {
  return guard([&] {
    GRE_CHECK(g && srcs, GRE_DEV_NULL, "null argument");
    GRE_CHECK(nsrc >= 1 && nsrc <= 64, GRE_DEV_NULL, "nsrc must be in [1, 64]");
    for (int i = 0; i < nsrc; ++i) GRE_CHECK(srcs[i] < g->nV, GRE_DEV_NULL, "source id >= nV");
    hipStream_t s = stream();
    const uint32_t nV = g->nV;
    // the first frontier: the distinct sources
    std::vector<uint32_t> uq(srcs, srcs + nsrc);
    std::sort(uq.begin(), uq.end());
    uq.erase(std::unique(uq.begin(), uq.end()), uq.end());
    // the pair is timed on the first call with a batch wide enough to show
    // the lanes' cost (a few sources finish their rows early); until then the
    // first candidate runs
    if (!g->ms_long && nsrc >= kTuneMinSources) ms_tune(g, srcs, nsrc, uq);
    DevBuf lev;
    if (level_out) {
      lev.alloc(4 * (size_t)nV * nsrc);
      hipLaunchKernelGGL(k_fill<int32_t>, dim3(gcap((uint64_t)nV * nsrc)), dim3(kB), 0, s, lev.as<int32_t>(),
                         (uint64_t)nV * nsrc, -1);
      GRE_HIP(hipGetLastError());
    }
    ms_run(g, srcs, nsrc, uq, level_out ? lev.as<int32_t>() : nullptr, g->ms_long ? g->ms_long : kMsCand[0][0],
           g->ms_long ? g->ms_chunk : kMsCand[0][1]);
    MsBufs& mb = ms_bufs(nV);
    if (edges_out) {
      GRE_HIP(hipMemsetAsync(mb.flag.p, 0, 8, s));
      hipLaunchKernelGGL(k_ms_edges, dim3(gcap(nV)), dim3(kB), 0, s, (const unsigned long long*)mb.vis.p,
                         g->off.as<uint64_t>(), nV, (unsigned long long*)mb.flag.p);
      GRE_HIP(hipGetLastError());
      GRE_HIP(hipMemcpyAsync(edges_out, mb.flag.p, 8, hipMemcpyDeviceToHost, s));
    }
    if (level_out)
      GRE_HIP(hipMemcpyAsync(level_out, lev.p, 4 * (size_t)nV * nsrc, hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
  });
}

int gre_bc(gre_graph* g, const uint32_t* srcs, int nsrc, double* bc_out, double* sigma_out, double* delta_out,
           int32_t* level_out) {
  return guard([&] {
    GRE_CHECK(g && srcs && nsrc >= 1, GRE_DEV_NULL, "bad argument");
    for (int i = 0; i < nsrc; ++i) GRE_CHECK(srcs[i] < g->nV, GRE_DEV_NULL, "source id >= nV");
    hipStream_t s = stream();
    const uint32_t nV = g->nV;
    if (!g->bc_flong && nsrc >= kTuneMinSources) bc_tune(g, srcs, nsrc);
    // per-source outputs: filled on the device, chunk by chunk, copied once at the end
    DevBuf xs, xd, xl;
    if (sigma_out) xs.alloc(8 * (size_t)nV * nsrc);
    if (delta_out) xd.alloc(8 * (size_t)nV * nsrc);
    if (level_out) xl.alloc(4 * (size_t)nV * nsrc);
    bc_run(g, srcs, nsrc, sigma_out ? xs.as<double>() : nullptr, delta_out ? xd.as<double>() : nullptr,
           level_out ? xl.as<int32_t>() : nullptr, g->bc_flong ? g->bc_flong : kBcCand[0][0],
           g->bc_flong ? g->bc_fchunk : kBcCand[0][1]);
    if (sigma_out) GRE_HIP(hipMemcpyAsync(sigma_out, xs.p, xs.bytes, hipMemcpyDeviceToHost, s));
    if (delta_out) GRE_HIP(hipMemcpyAsync(delta_out, xd.p, xd.bytes, hipMemcpyDeviceToHost, s));
    if (level_out) GRE_HIP(hipMemcpyAsync(level_out, xl.p, xl.bytes, hipMemcpyDeviceToHost, s));
    if (bc_out) GRE_HIP(hipMemcpyAsync(bc_out, bc_bufs(nV, 1, 1).bc.p, 8 * (size_t)nV, hipMemcpyDeviceToHost, s));
    GRE_HIP(hipStreamSynchronize(s));
  });
}

int gre_traversal_set_chunks(gre_graph* g, int which, uint64_t long_rows, uint64_t chunk) {
  return guard([&] {
    GRE_CHECK(g, GRE_DEV_NULL, "null graph");
    GRE_CHECK(which == GRE_CHUNKS_MSBFS || which == GRE_CHUNKS_BC_FORWARD, GRE_DEV_NULL, "unknown traversal pass");
    GRE_CHECK(long_rows == 0 || (long_rows < (1ull << 31) && chunk >= 1 && chunk < (1ull << 32)), GRE_DEV_NULL,
              "long_rows must be below 2^31 and chunk in [1, 2^32)");
    uint64_t* p = which == GRE_CHUNKS_MSBFS ? &g->ms_long : &g->bc_flong;
    uint64_t* q = which == GRE_CHUNKS_MSBFS ? &g->ms_chunk : &g->bc_fchunk;
    *p = long_rows;
    *q = long_rows ? chunk : 0;
  });
}

int gre_traversal_get_chunks(gre_graph* g, int which, uint64_t* long_rows, uint64_t* chunk) {
  return guard([&] {
    GRE_CHECK(g && long_rows && chunk, GRE_DEV_NULL, "null argument");
    GRE_CHECK(which == GRE_CHUNKS_MSBFS || which == GRE_CHUNKS_BC_FORWARD, GRE_DEV_NULL, "unknown traversal pass");
    *long_rows = which == GRE_CHUNKS_MSBFS ? g->ms_long : g->bc_flong;
    *chunk = which == GRE_CHUNKS_MSBFS ? g->ms_chunk : g->bc_fchunk;
  });
}

}  // extern "C"
