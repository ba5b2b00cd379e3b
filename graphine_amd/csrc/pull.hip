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

void PullBuffers::make(PullPlan& p, gre_graph* g, bool csc, int ipt, size_t tbytes) {
  const uint32_t te = (uint32_t)(kPullBlock * ipt);
  p.off = (csc ? g->coff : g->off).as<uint64_t>();
  p.idx = (csc ? g->cidx : g->idx).as<uint32_t>();
  p.tile_rb = graph_tiles(g, csc, te, &p.ntiles);
  p.E = g->nE;
  p.V = g->nV;
  p.ipt = ipt;
  carryP.alloc(tbytes * p.ntiles);
  carryO.alloc(tbytes * p.ntiles);
  part.alloc(2 * sizeof(double) * p.ntiles);
  const uint64_t nfb = (p.ntiles + kPullBlock - 1) / kPullBlock;
  part2.alloc(2 * sizeof(double) * nfb);
  p.carryP = carryP.p;
  p.carryO = carryO.p;
  p.part = part.as<double>();
  p.part2 = part2.as<double>();
}

}  // namespace gre
