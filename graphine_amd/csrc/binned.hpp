// Binned (propagation-blocking) forward pull of the single-vector FP32
// PageRank: per-handle buffers and one iteration.  The layout is BinPlan
// (graph.hpp), built by graph_binpart (binned.hip).
#pragma once
#include "graph.hpp"

namespace gre {

// Per-handle working memory of the binned forward.
struct BinBuffers {
  DevBuf binval;           // float [Eb] each slot's group sum, bin order
  DevBuf gacc;             // int64 [nmrows] sums of the bins cut into several slices (zero between iterations)
  DevBuf part, part2, part3;
  unsigned fin_grid = 0;   // workgroups of the finishing pass
  uint32_t tail_blocks = 0;  // accumulate workgroups after the slices, for the rows without in-edges
  const XPart* for_xp = nullptr;  // the layout these were made for
  void make(const XPart& xp, uint32_t nV, hipStream_t s);
};

// One iteration's operands.  Ranks, personalisation and out-degrees are in the
// layout's row order; contributions are indexed by row.
struct BinIter {
  // weighted: the normalised edge weights in word order (BinPlan::wpos), and
  // contributions are the ranks themselves; nullptr: contributions are rank /
  // out-degree
  const float* wbin = nullptr;
  const float* contrib_in = nullptr;
  float* contrib_out = nullptr;
  float* x = nullptr;            // ranks, updated in place
  const float* p_r = nullptr;    // personalisation in row order, or nullptr (uniform)
  const int32_t* outdeg_r = nullptr;
  const double* scal_in = nullptr;  // [0] = dangling mass of the ranks read
  double* scal_out = nullptr;       // (dangling mass, L1 change) of the ranks written
  double damp = 0.85, invn = 0.0;
  bool resid = true;               // accumulate the L1 change
};

void bin_forward_iter(const XPart& xp, const BinBuffers& bb, const BinIter& a, uint32_t nV, hipStream_t s);

// One iteration of the single-vector FP32 adjoint over the binned layout of the
// reversed graph (graph_reversed): rows are the reversed graph's, every vector
// in its row order, lambda's scatter reads lin itself.  Buffers as the forward's
// (BinBuffers made for that layout).
struct BinAdjIter {
  const float* wbin = nullptr;   // weighted: normalised weights in word order; nullptr: unweighted
  const float* lin = nullptr;    // lambda read
  float* lout = nullptr;         // lambda written
  float* wl = nullptr;           // (W lambda) written
  const float* c_r = nullptr;
  const float* p_r = nullptr;    // personalisation, or nullptr (uniform)
  const int32_t* deg_r = nullptr;  // out-degree (of the graph, not the reversed one)
  const double* scal_in = nullptr;  // [0] = p.lambda of lin
  double* scal_out = nullptr;       // (p.lambda, L1 change) of lout
  const uint32_t* mx_in = nullptr;  // bits of max |lin|
  uint32_t* mx_out = nullptr;       // max |lout| is folded in (zero it first)
  double F = 1.0;                   // bound on sum_v |w_uv| over every row u
  double damp = 0.85, invn = 0.0;
};
void bin_adjoint_iter(const XPart& xp, const BinBuffers& bb, const BinAdjIter& a, uint32_t nV, hipStream_t s);
// *mx = max(*mx, bits of max |x[i]|); max over uint32 values; zero fill
void bin_absmax(const float* x, uint64_t n, uint32_t* mx, hipStream_t s);
void bin_maxu(const uint32_t* x, uint64_t n, uint32_t* mx, hipStream_t s);
void bin_zero_u32(uint32_t* p, uint32_t n, hipStream_t s);

// The engine's FP64 sum pull over the binned layout of graph_binpart_engine
// (sources staged in compact source order, rows = the vertices with in-edges
// in vertex order): C[v] += the sum of S over v's in-neighbours, as exact
// two-word fixed-point sums, for every vertex v with in-edges.
struct EngBinBuffers {
  DevBuf binval;   // double [Eb] each slot's group sum, bin order
  DevBuf trash;    // the scatter's slots for lanes with nothing to write
  DevBuf gacc;     // int64 [2 nmrows] two-word sums of the bins cut into slices
  DevBuf Sc;       // double [nsrc] the scatter data in compact source order
  DevBuf smax;     // the largest |S| (bits of a double), which sets the fixed-point scales
  const XPart* for_xp = nullptr;
  void make(const XPart& xp, uint32_t nsrc, hipStream_t s);
};
// An update fused into the sum pull: for every vertex v with in-edges,
// X[v] = a + b (C[v] + sum), then C[v] = 0.  X = nullptr: C[v] += sum only.
struct EngFuse {
  double* X = nullptr;
  double a = 0.0, b = 0.0;
};
// fresh_s: S changed since the last call (its compact copy and largest value
// are remade)
void bin_engine_sum(const XPart& xp, const EngBinBuffers& bb, const double* S, const uint32_t* cid, uint32_t nV,
                    bool fresh_s, double* C, const EngFuse& f, hipStream_t s);

// Fixed-point scale of the accumulate pass: contributions are added as
// round(c * 2^kBinFix) in 64-bit integers, so every sum is exact and
// independent of the order of the additions.  |sum| must stay below
// 2^(63 - kBinFix) = 32.  A row's sum is at most the L1 norm of the ranks
// (each source spreads rank / out-degree over its out-edges), and with a
// personalisation of L1 norm at most 1 the iteration keeps that norm at most
// 1; the schedule refuses personalisations above kBinMaxPL1.
constexpr int kBinFix = 58;
constexpr double kBinMaxPL1 = 1.001;

}  // namespace gre
