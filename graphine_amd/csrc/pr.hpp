// The PageRank handle shared by the single-vector and batched kernels.
#pragma once
#include "binned.hpp"
#include "pull.hpp"

#include <cstring>
#include <vector>

namespace gre {
constexpr int kMaxBatch = 64;
struct BatchState {
  DevBuf carryP, carryO, part, scal;
  uint64_t ntiles = 0;
  int scal_cap = 0;
  // schedule of each direction (0 forward over CSC, 1 adjoint over CSR):
  // 0 = not chosen yet, GRE_SCHED_TILE = tile walk, GRE_SCHED_ROWS = group per row
  int sched[2] = {0, 0};
  // fixed-iteration runs replayed from captured hipGraphs, per direction
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  std::vector<uint64_t> key[2];
  int post_cur[2] = {0, 0};
  void drop(int d) {
    if (exec[d]) { (void)hipGraphExecDestroy(exec[d]); exec[d] = nullptr; }
    if (graph[d]) { (void)hipGraphDestroy(graph[d]); graph[d] = nullptr; }
    key[d].clear();
  }
  BatchState() = default;
  BatchState(const BatchState&) = delete;
  BatchState& operator=(const BatchState&) = delete;
  ~BatchState() { drop(0); drop(1); }
};
}  // namespace gre

struct gre_pr {
  gre_graph* g = nullptr;
  int dtype = GRE_F32;
  int k = 1;
  int weighted = 0;
  int ipt = 0;            // 0 = autotune on first run
  int schedule = 0;       // 0 = autotune; else GRE_SCHED_* (gre_dev.h)
  bool nt = false;        // non-temporal neighbour-id stream
  bool compact = false;   // hub-sorted compact source ids (k == 1)
  bool part = false;      // source-range partitioned pull (implies compact)
  bool bin = false;       // binned pull (single vector, FP32, unweighted)
  bool tuning = false;    // the autotuner is running: keep the candidates' partitions cached
  int ranges = 64;        // tail source ranges of the partitioned pull
  uint32_t shared = 1u << 18;  // hottest sources processed as one range by all XCDs
  std::shared_ptr<gre::XPart> xp;  // the partition in use
  gre::DevBuf xsum;       // per-segment partial sums (compute type)
  gre::DevBuf wpart;      // weighted: normalised weights in the partition's entry order
  std::shared_ptr<gre::XPart> wpart_xp;  // the partition and weights version wpart was made for
  uint64_t wpart_ver = ~0ULL;
  uint64_t wver = 0;      // bumped by every gre_pr_set_edge_weights
  gre::PullPlan plan_x, plan_px;   // segment pull; per-row pull over the segment sums
  gre::PullBuffers buf_x, buf_px;
  int plan_ipt_x = 0;
  bool has_p = false;
  double p_l1 = 1.0;      // L1 norm of the personalisation (1 when uniform)
  gre::BinBuffers bbuf;   // binned pull: bin array, multi-slice sums, partials
  // forward state
  gre::DevBuf x[2];       // ranks (weighted: ping-pong) ; unweighted: x[0] in place
  gre::DevBuf contrib[2]; // gathered per-source values, ping-pong: x/outdeg (unweighted), or the
                          // ranks in compact source order (weighted, compact and part schedules)
  int xcur = 0;           // x[xcur] holds the current ranks (always 0 when unweighted, k == 1)
  int ccur = 0;           // contrib[ccur] holds the current x/outdeg (unweighted, k == 1)
  gre::DevBuf p;          // personalisation (nV*k), optional
  gre::DevBuf p_r;        // the personalisation in the partition's row order (relabelled rows)
  uint64_t pver = 0;      // bumped by every set_personalization
  uint64_t p_r_ver = ~0ULL;                 // the p version p_r holds ...
  std::shared_ptr<gre::XPart> p_r_xp;       // ... in this partition's row order
  // A run over relabelled partition rows leaves its ranks in x[xcur] in the
  // partition's row order (rows_xp: that partition); whatever reads the ranks
  // first turns them back into vertex order.
  bool x_rows = false;
  std::shared_ptr<gre::XPart> rows_xp;
  gre::DevBuf wraw, wcsr, wcsc, rowsum;  // weights (CSR raw, CSR norm, CSC norm), FP64 row sums
  // adjoint
  gre::DevBuf lam[2], c, wl;
  gre::DevBuf lc[2];      // adjoint, compact schedules: lambda in compact target order, ping-pong
  // adjoint, part schedules: the same partition over the CSR (rows = sources,
  // ranges of compact targets), its segment sums and weights in its entry order
  std::shared_ptr<gre::XPart> xpa;
  gre::DevBuf asum, wparta;
  std::shared_ptr<gre::XPart> wparta_xp;
  uint64_t wparta_ver = ~0ULL;
  gre::PullPlan plan_xa, plan_pxa;
  gre::PullBuffers buf_xa, buf_pxa;
  int plan_ipt_xa = 0;
  // adjoint, binned schedule: the binned layout of the reversed graph, its
  // buffers, c, p, out-degrees and (W lambda) in its row order, the weights in
  // its word order, max |lambda| per iteration, and the bound of the row sums
  std::shared_ptr<gre::XPart> xpb;
  gre::BinBuffers bbufb;
  gre::DevBuf cb_r, pb_r, degb_r, wlb_r, wbinb, mxb;
  std::shared_ptr<gre::XPart> wbinb_xp;
  uint64_t wbinb_ver = ~0ULL;
  double maxdeg_b = 0.0, wmax_b = 0.0;
  int lcur = 0;
  bool have_adj = false;
  bool have_fwd = false;
  gre::DevBuf scal;       // [2*(maxit+1)] per-iteration (acc0, acc1)
  int scal_cap = 0;
  // plans
  gre::PullPlan plan_csc, plan_csr;
  gre::PullBuffers buf_csc, buf_csr;
  int plan_ipt_csc = 0, plan_ipt_csr = 0;
  gre::BatchState bstate;
  // fixed-iteration forward runs replayed from a captured hipGraph
  bool graphs = true;
  hipGraph_t fwd_graph = nullptr;
  hipGraphExec_t fwd_exec = nullptr;
  std::vector<uint64_t> fwd_key;   // what the capture baked in (sizes, pointers, parameters)
  int fwd_post_xcur = 0, fwd_post_ccur = 0;
  ~gre_pr() {
    if (fwd_exec) (void)hipGraphExecDestroy(fwd_exec);
    if (fwd_graph) (void)hipGraphDestroy(fwd_graph);
  }
};

namespace gre {
double batched_forward(gre_pr* P, double damp, int iters, double tol, int* done);
double batched_adjoint(gre_pr* P, double damp, int iters, double tol, int* done);
void batched_targets(gre_pr* P, const uint32_t* targets_dev, int64_t nt);
void batched_gradients(gre_pr* P, double damp, float* gW, float* graw, void* gp);
}  // namespace gre
