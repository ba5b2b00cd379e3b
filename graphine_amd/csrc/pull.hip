// Host side of the tile-segmented pull and the pair reduction.
#include "pull.hpp"

namespace gre {

__global__ void k_reduce_pairs(const double* __restrict__ a, uint64_t na, const double* __restrict__ b,
                               uint64_t nb, double* __restrict__ out) {
  __shared__ double sh0[1024], sh1[1024];
  double s0 = 0.0, s1 = 0.0;
  for (uint64_t i = threadIdx.x; i < na; i += blockDim.x) { s0 += a[2 * i]; s1 += a[2 * i + 1]; }
  for (uint64_t i = threadIdx.x; i < nb; i += blockDim.x) { s0 += b[2 * i]; s1 += b[2 * i + 1]; }
  sh0[threadIdx.x] = s0;
  sh1[threadIdx.x] = s1;
  __syncthreads();
  for (unsigned o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sh0[threadIdx.x] += sh0[threadIdx.x + o];
      sh1[threadIdx.x] += sh1[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[0] = sh0[0]; out[1] = sh1[0]; }
}

__global__ __launch_bounds__(256) void k_reduce_pairs_part(const double* __restrict__ a, uint64_t na,
                                                           const double* __restrict__ b, uint64_t nb,
                                                           double* __restrict__ out) {
  __shared__ double red[8];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  double s0 = 0.0, s1 = 0.0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < na; i += stride) {
    const double2 v = reinterpret_cast<const double2*>(a)[i];
    s0 += v.x;
    s1 += v.y;
  }
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nb; i += stride) {
    const double2 v = reinterpret_cast<const double2*>(b)[i];
    s0 += v.x;
    s1 += v.y;
  }
  block_sum2(s0, s1, red);
  if (threadIdx.x == 0) { out[2 * blockIdx.x] = s0; out[2 * blockIdx.x + 1] = s1; }
}

void reduce_pairs(const double* a, uint64_t na, const double* b, uint64_t nb, double* out, double* tmp,
                  hipStream_t s) {
  if (na + nb <= 16384) {
    hipLaunchKernelGGL(k_reduce_pairs, dim3(1), dim3(1024), 0, s, a, na, b, nb, out);
    return;
  }
  hipLaunchKernelGGL(k_reduce_pairs_part, dim3(kRedBlocks), dim3(256), 0, s, a, na, b, nb, tmp);
  hipLaunchKernelGGL(k_reduce_pairs, dim3(1), dim3(1024), 0, s, tmp, (uint64_t)kRedBlocks, tmp, (uint64_t)0, out);
}

void PullBuffers::make(PullPlan& p, gre_graph* g, bool csc, int ipt, size_t tbytes) {
  const uint32_t te = (uint32_t)(kPullBlock * ipt);
  p.off = (csc ? g->coff : g->off).as<uint64_t>();
  p.off32 = nullptr;
  p.idx = (csc ? g->cidx : g->idx).as<uint32_t>();
  p.tile_rb = graph_tiles(g, csc, te, &p.ntiles);
  p.E = g->nE;
  p.V = g->nV;
  p.ipt = ipt;
  p.xmap = nullptr;
  p.slot = nullptr;
  p.grid = 0;
  alloc_for(p, tbytes);
}

void PullBuffers::make(PullPlan& p, XPart& xp, int ipt, size_t tbytes) {
  const uint32_t te = (uint32_t)(kPullBlock * ipt);
  const XPart::Tiles& T = xpart_tiles(xp, te, stream());
  p.off = nullptr;
  p.off32 = xp.seg_off.as<uint32_t>();
  p.idx = xp.idx.as<uint32_t>();
  p.tile_rb = T.rb.as<uint32_t>();
  p.ntiles = T.ntiles;
  p.E = xp.class_start[8];
  p.V = (uint32_t)xp.nseg;
  p.ipt = ipt;
  p.xmap = T.xmap.as<uint32_t>();
  p.slot = nullptr;
  p.grid = T.grid;
  alloc_for(p, tbytes);
}

void PullBuffers::make_pairs(PullPlan& p, XPart& xp, uint32_t nV, int ipt, size_t tbytes) {
  const uint32_t te = (uint32_t)(kPullBlock * ipt);
  const XPart::PairTiles& T = xpart_pair_tiles(xp, nV, te, stream());
  p.off = nullptr;
  p.off32 = xp.pair_off.as<uint32_t>();
  p.idx = xp.pair_list.as<uint32_t>();
  p.slot = T.slot.as<uint16_t>();
  p.words = T.words.as<uint2>();
  p.base = T.base.as<uint32_t>();
  p.K = xp.K;
  p.tile_rb = T.rb.as<uint32_t>();
  p.ntiles = T.ntiles;
  p.E = xp.nseg;
  p.V = nV;
  p.ipt = ipt;
  p.xmap = nullptr;
  p.grid = 0;
  alloc_for(p, tbytes);
}

void PullBuffers::alloc_for(PullPlan& p, size_t tbytes) {
  carryP.alloc(tbytes * p.ntiles);
  carryO.alloc(tbytes * p.ntiles);
  part.alloc(2 * sizeof(double) * p.ntiles);
  const uint64_t nfb = (p.ntiles + kPullBlock - 1) / kPullBlock;
  part2.alloc(2 * sizeof(double) * nfb);
  part3.alloc(2 * sizeof(double) * kRedBlocks);
  p.carryP = carryP.p;
  p.carryO = carryO.p;
  p.part = part.as<double>();
  p.part2 = part2.as<double>();
  p.part3 = part3.as<double>();
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr int kProbeU = 8;   // 16-byte loads in flight per thread

// Streaming copy: each 256-thread block moves 256 * kProbeU 16-byte vectors,
// all loads issued before the first store, stores non-temporal.  One block per
// 32 KB, so a 4 GiB copy launches 131,072 workgroups.  The calibration kernel
// for HBM byte counters and the achievable copy bandwidth.
__global__ __launch_bounds__(256) void k_copy_probe(const u32x4* __restrict__ in, u32x4* __restrict__ out,
                                                    uint64_t n) {
  const uint64_t base = (uint64_t)blockIdx.x * 256 * kProbeU + threadIdx.x;
  u32x4 v[kProbeU];
  if (base + (uint64_t)(kProbeU - 1) * 256 < n) {
#pragma unroll
    for (int j = 0; j < kProbeU; ++j) v[j] = in[base + (uint64_t)j * 256];
#pragma unroll
    for (int j = 0; j < kProbeU; ++j) __builtin_nontemporal_store(v[j], out + base + (uint64_t)j * 256);
  } else {
    for (int j = 0; j < kProbeU; ++j)
      if (base + (uint64_t)j * 256 < n) out[base + (uint64_t)j * 256] = in[base + (uint64_t)j * 256];
  }
}

// Streaming read: the same access pattern without the stores; each thread
// folds its vectors into one word so no load is dead.  The achievable rate of
// a read-only stream (the PageRank iteration reads about 15 bytes per byte it
// writes).
__global__ __launch_bounds__(256) void k_read_probe(const u32x4* __restrict__ in, uint64_t n,
                                                    unsigned* __restrict__ out) {
  const uint64_t base = (uint64_t)blockIdx.x * 256 * kProbeU + threadIdx.x;
  u32x4 v[kProbeU];
  unsigned acc = 0;
  if (base + (uint64_t)(kProbeU - 1) * 256 < n) {
#pragma unroll
    for (int j = 0; j < kProbeU; ++j) v[j] = in[base + (uint64_t)j * 256];
#pragma unroll
    for (int j = 0; j < kProbeU; ++j) acc ^= v[j].x ^ v[j].y ^ v[j].z ^ v[j].w;
  } else {
    for (int j = 0; j < kProbeU; ++j)
      if (base + (uint64_t)j * 256 < n) {
        const u32x4 w = in[base + (uint64_t)j * 256];
        acc ^= w.x ^ w.y ^ w.z ^ w.w;
      }
  }
  if (acc == 0x9e3779b9u) out[blockIdx.x] = acc;   // practically never: keeps the loads live
}

// Gather probe: ids drawn uniformly from [0, n_tab) by a hash, then one
// 4-byte gather each from a table of n_tab floats (ids streamed, 8 per thread)
__global__ __launch_bounds__(256) void k_gather_ids(uint32_t* __restrict__ ids, uint64_t n, uint64_t n_tab) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t z = (i + 1) * GRE_K0;
    z = (z ^ (z >> 30)) * GRE_K1;
    z = (z ^ (z >> 27)) * GRE_K2;
    ids[i] = (uint32_t)((z ^ (z >> 31)) % n_tab);
  }
}

__global__ __launch_bounds__(256) void k_gather_probe(const uint32_t* __restrict__ ids, const float* __restrict__ tab,
                                                      uint64_t n, float* __restrict__ out) {
  const uint64_t base = (uint64_t)blockIdx.x * 256 * 8;
  uint32_t u[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint64_t i = base + j * 256 + threadIdx.x;
    u[j] = i < n ? __builtin_nontemporal_load(ids + i) : 0u;
  }
  // unconditional gathers (u = 0 past the end), all in flight before the sum
  float v[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = tab[u[j]];
  float acc = 0.f;
#pragma unroll
  for (int j = 0; j < 8; ++j) acc += base + j * 256 + threadIdx.x < n ? v[j] : 0.f;
  out[(uint64_t)blockIdx.x * 256 + threadIdx.x] = acc;
}

}  // namespace gre

extern "C" int gre_gather_probe(uint64_t table_entries, uint64_t gathers, int reps, double* ms_per_pass) {
  using namespace gre;
  return guard([&] {
    GRE_CHECK(ms_per_pass && reps > 0 && table_entries > 0 && table_entries <= 0xffffffffULL && gathers >= 2048,
              GRE_DEV_NULL, "bad argument");
    hipStream_t s = stream();
    const uint64_t nblk = (gathers + 2047) / 2048;
    GRE_CHECK(nblk <= 0x7fffffffULL, GRE_DEV_NULL, "too many gathers");
    DevBuf ids, tab, out;
    ids.alloc(4 * gathers);
    tab.alloc(4 * table_entries);
    out.alloc(4 * 256 * nblk);
    GRE_HIP(hipMemsetAsync(tab.p, 0, 4 * table_entries, s));
    hipLaunchKernelGGL(k_gather_ids, dim3(grid_for(gathers, 256) < 65536u ? grid_for(gathers, 256) : 65536u),
                       dim3(256), 0, s, ids.as<uint32_t>(), gathers, table_entries);
    hipLaunchKernelGGL(k_gather_probe, dim3((unsigned)nblk), dim3(256), 0, s, ids.as<uint32_t>(), tab.as<float>(),
                       gathers, out.as<float>());
    hipEvent_t e0, e1;
    GRE_HIP(hipEventCreate(&e0));
    GRE_HIP(hipEventCreate(&e1));
    GRE_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) {
      ProfScope ps("gather_probe");
      hipLaunchKernelGGL(k_gather_probe, dim3((unsigned)nblk), dim3(256), 0, s, ids.as<uint32_t>(), tab.as<float>(),
                         gathers, out.as<float>());
    }
    GRE_HIP(hipEventRecord(e1, s));
    GRE_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    GRE_HIP(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_pass = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  });
}

// Times `reps` launches of a streaming probe over `bytes` bytes (copy: in -> out;
// read: in only) on the library stream.
static int stream_probe(uint64_t bytes, int reps, bool copy, double* ms_per_pass) {
  using namespace gre;
  return guard([&] {
    GRE_CHECK(ms_per_pass && reps > 0 && bytes >= 16, GRE_DEV_NULL, "bad argument");
    hipStream_t s = stream();
    const uint64_t n = bytes / 16;
    const uint64_t nblk = (n + 256 * kProbeU - 1) / (256 * kProbeU);
    GRE_CHECK(nblk <= 0x7fffffffULL, GRE_DEV_NULL, "probe too large");
    DevBuf a, b;
    a.alloc(16 * n);
    if (copy) b.alloc(16 * n);
    else b.alloc(4 * nblk);
    GRE_HIP(hipMemsetAsync(a.p, 1, 16 * n, s));
    hipEvent_t e0, e1;
    GRE_HIP(hipEventCreate(&e0));
    GRE_HIP(hipEventCreate(&e1));
    auto launch = [&] {
      if (copy)
        hipLaunchKernelGGL(k_copy_probe, dim3((unsigned)nblk), dim3(256), 0, s, a.as<u32x4>(), b.as<u32x4>(), n);
      else
        hipLaunchKernelGGL(k_read_probe, dim3((unsigned)nblk), dim3(256), 0, s, a.as<u32x4>(), n, b.as<unsigned>());
    };
    launch();
    GRE_HIP(hipGetLastError());
    GRE_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) {
      ProfScope ps(copy ? "copy_probe" : "read_probe");
      launch();
    }
    GRE_HIP(hipEventRecord(e1, s));
    GRE_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    GRE_HIP(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_pass = ms / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  });
}

extern "C" int gre_copy_probe(uint64_t bytes, int reps, double* ms_per_copy) {
  return stream_probe(bytes, reps, true, ms_per_copy);
}

extern "C" int gre_read_probe(uint64_t bytes, int reps, double* ms_per_read) {
  return stream_probe(bytes, reps, false, ms_per_read);
}
