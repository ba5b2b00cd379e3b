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
