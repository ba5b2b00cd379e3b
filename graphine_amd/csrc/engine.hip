// GRE's BSP superstep on the device: SynchronousEngine<VC>::run<dir>()
// (include/Engine/Engine.hpp:180-406 master, :422-814 workers) at one process.
//
// Superstep k, as GRE runs it once the two MPI sync phases are no-ops:
//   compute  every vertex u in the current frontier queue scatters along its
//            Out / In / InOut edges; each message is combined into the
//            destination's combine data, and a true combine() marks the
//            destination in nextSet (Engine.hpp:592-625, :1385-1389);
//   halt     assert_to_halt(u) clears u from the currSet bitmap (:617-621);
//   apply    every vertex marked in nextSet runs apply(); a true result sets it
//            in currSet (:757-775);
//   rebuild  the frontier queue is rebuilt from the whole currSet bitmap, which
//            persists across supersteps (:777-798);
//   vote     the active count goes back to the host, which stops at zero or at
//            max_nSupersteps (:386-400).
// Here: compute = one edge-balanced push kernel (frontier degrees are
// prefix-summed so every thread gets the same number of edges, and a hub's
// edges spread over many waves); combine = one atomic per message (atomicAdd,
// atomicMin, first-writer CAS), which plays the role of GRE's 8192 striped
// spin-locks; apply + halt + count are one thread-per-vertex kernel that
// rebuilds the bitmap words with wave ballots; rebuild = block scan + write.
#include "binned.hpp"
#include "graph.hpp"
#include "pull.hpp"

#include <algorithm>
#include <memory>
#include <chrono>
#include <type_traits>

namespace gre {
namespace {

constexpr int kB = 256;
constexpr int kVB = 4096;           // vertices per block in apply/count/write
constexpr int kEB = kB * 8;         // edges per block in the push expand

struct EngCtx {
  void* S;
  void* C;
  void* X;
  const void* aux;
  int32_t level;
  const int32_t* outdeg;
  const uint32_t* w_csr;            // SSSP weights, CSR order
};

struct VtxState { int32_t level; int32_t numPaths; };  // BC.hpp:12-17

// ---- vertex programs --------------------------------------------------------
// init / scatter+combine / apply restate include/Kernels/*.hpp; scatter+combine
// returns combine()'s result for the message of edge (u -> v).
template <int P> struct Prog;

template <> struct Prog<GRE_PROG_PAGERANK> {            // PageRank.hpp:18-126
  static constexpr bool kHalt = false;                  // assert_to_halt -> false
  __device__ static bool init(const EngCtx& c, uint32_t v) {
    double* pr = (double*)c.X; double* ps = (double*)c.S;
    pr[v] = 0.15;
    const int32_t d = c.outdeg[v];
    if (d > 0) { ps[v] = pr[v] / (double)d; return true; }
    return false;
  }
  __device__ static bool combine(const EngCtx& c, uint32_t u, uint32_t v, uint64_t) {
    atomic_add_f64((double*)c.C + v, ((const double*)c.S)[u]);
    return true;
  }
  static constexpr double kResetProb = 0.15;
  __device__ static bool apply(const EngCtx& c, uint32_t v) {
    double* pr = (double*)c.X; double* cm = (double*)c.C;
    pr[v] = kResetProb + (1.0 - kResetProb) * cm[v];
    cm[v] = 0.0;
    return false;
  }
};

template <> struct Prog<GRE_PROG_PAGERANK_DELTA> {      // PageRank.hpp:129-234
  static constexpr bool kHalt = true;
  __device__ static bool init(const EngCtx& c, uint32_t v) { return Prog<GRE_PROG_PAGERANK>::init(c, v); }
  __device__ static bool combine(const EngCtx& c, uint32_t u, uint32_t v, uint64_t e) {
    return Prog<GRE_PROG_PAGERANK>::combine(c, u, v, e);
  }
  __device__ static bool apply(const EngCtx& c, uint32_t v) {
    double* pr = (double*)c.X; double* cm = (double*)c.C; double* ps = (double*)c.S;
    const double delta = (1.0 - 0.15) * cm[v];
    pr[v] += delta;
    cm[v] = 0.0;
    const int32_t d = c.outdeg[v];
    if (delta > 1e-5 && d > 0) { ps[v] = delta / (double)d; return true; }
    return false;
  }
};

template <> struct Prog<GRE_PROG_BFS> {                 // BFS.hpp:18-86
  static constexpr bool kHalt = true;
  __device__ static bool init(const EngCtx& c, uint32_t v) {
    uint64_t* pred = (uint64_t*)c.X; uint64_t* ps = (uint64_t*)c.S;
    pred[v] = v;   // get_vid(lv) == lv at one process
    ps[v] = pred[v];
    return true;
  }
  __device__ static bool combine(const EngCtx& c, uint32_t u, uint32_t v, uint64_t) {
    unsigned long long* cm = (unsigned long long*)c.C;
    if (cm[v] != ~0ULL) return false;
    const unsigned long long m = ((const unsigned long long*)c.S)[u];
    return atomicCAS(cm + v, ~0ULL, m) == ~0ULL;
  }
  __device__ static bool apply(const EngCtx& c, uint32_t v) {
    uint64_t* pred = (uint64_t*)c.X; uint64_t* cm = (uint64_t*)c.C; uint64_t* ps = (uint64_t*)c.S;
    if (pred[v] == ~0ULL) { pred[v] = cm[v]; ps[v] = pred[v]; return true; }
    return false;
  }
};

template <> struct Prog<GRE_PROG_CC> {                  // CC.hpp:18-125
  static constexpr bool kHalt = true;
  __device__ static bool init(const EngCtx& c, uint32_t v) {
    uint64_t* cc = (uint64_t*)c.X; uint64_t* cs = (uint64_t*)c.S; uint64_t* cm = (uint64_t*)c.C;
    cc[v] = v; cs[v] = cc[v]; cm[v] = cc[v];
    return true;
  }
  __device__ static bool combine(const EngCtx& c, uint32_t u, uint32_t v, uint64_t) {
    unsigned long long* cm = (unsigned long long*)c.C;
    const unsigned long long m = ((const unsigned long long*)c.S)[u];
    if (!(cm[v] > m)) return false;
    return m < atomicMin(cm + v, m);
  }
  __device__ static bool apply(const EngCtx& c, uint32_t v) {
    uint64_t* cc = (uint64_t*)c.X; uint64_t* cm = (uint64_t*)c.C; uint64_t* cs = (uint64_t*)c.S;
    if (cm[v] >= cc[v]) return false;
    cc[v] = cm[v]; cs[v] = cc[v];
    return true;
  }
};

template <> struct Prog<GRE_PROG_SSSP> {                // SSSP.hpp:19-96 (states = C = X)
  static constexpr bool kHalt = true;
  __device__ static bool init(const EngCtx& c, uint32_t v) {
    uint32_t* st = (uint32_t*)c.C; uint32_t* last = (uint32_t*)c.S;
    st[v] = 0; last[v] = st[v];
    return true;
  }
  __device__ static bool combine(const EngCtx& c, uint32_t u, uint32_t v, uint64_t epos) {
    uint32_t* st = (uint32_t*)c.C;
    const uint32_t d = ((const uint32_t*)c.S)[u] + c.w_csr[epos];
    if (!(d < st[v])) return false;
    return d < atomicMin(st + v, d);
  }
  __device__ static bool apply(const EngCtx& c, uint32_t v) {
    ((uint32_t*)c.S)[v] = ((uint32_t*)c.C)[v];
    return true;
  }
};

template <> struct Prog<GRE_PROG_BC_FORWARD> {          // BC.hpp:22-95
  static constexpr bool kHalt = true;
  __device__ static bool init(const EngCtx& c, uint32_t v) {
    VtxState* s = (VtxState*)c.X; int32_t* np = (int32_t*)c.S;
    s[v].numPaths = 1; s[v].level = 0; np[v] = s[v].numPaths;
    return true;
  }
  __device__ static bool combine(const EngCtx& c, uint32_t u, uint32_t v, uint64_t) {
    atomicAdd((int32_t*)c.C + v, ((const int32_t*)c.S)[u]);   // int32, wraps like GRE's int
    return true;
  }
  __device__ static bool apply(const EngCtx& c, uint32_t v) {
    VtxState* s = (VtxState*)c.X; int32_t* sum = (int32_t*)c.C; int32_t* np = (int32_t*)c.S;
    if (s[v].level < 0) { s[v].level = c.level; s[v].numPaths = sum[v]; np[v] = s[v].numPaths; return true; }
    return false;
  }
};

template <> struct Prog<GRE_PROG_BC_BACKWARD> {         // BC.hpp:97-168
  static constexpr bool kHalt = true;
  __device__ static bool init(const EngCtx& c, uint32_t v) {
    const VtxState* bs = (const VtxState*)c.aux; double* pw = (double*)c.S; double* dep = (double*)c.X;
    if (bs[v].level == c.level) { pw[v] = (1.0 + dep[v]) / (double)bs[v].numPaths; return true; }
    return false;
  }
  __device__ static bool combine(const EngCtx& c, uint32_t u, uint32_t v, uint64_t) {
    atomic_add_f64((double*)c.C + v, ((const double*)c.S)[u]);
    return true;
  }
  __device__ static bool apply(const EngCtx& c, uint32_t v) {
    const VtxState* bs = (const VtxState*)c.aux; double* sum = (double*)c.C; double* dep = (double*)c.X;
    if (bs[v].level == c.level - 1) dep[v] = sum[v] * (double)bs[v].numPaths;
    return false;
  }
};

template <> struct Prog<GRE_PROG_COUNTIN> {             // CountIn.hpp:18-106
  static constexpr bool kHalt = true;
  __device__ static bool init(const EngCtx& c, uint32_t v) {
    ((int32_t*)c.X)[v] = 0; ((int32_t*)c.S)[v] = 1;
    return true;
  }
  __device__ static bool combine(const EngCtx& c, uint32_t u, uint32_t v, uint64_t) {
    atomicAdd((int32_t*)c.C + v, ((const int32_t*)c.S)[u]);
    return true;
  }
  __device__ static bool apply(const EngCtx& c, uint32_t v) {
    ((int32_t*)c.X)[v] = ((int32_t*)c.C)[v];
    return false;
  }
};

// ---- pull form of the compute phase ---------------------------------------
// For a dense frontier every destination v reads the messages its active
// neighbours would send it and folds them with the program's combine algebra
// in registers, then merges the result into its combine data once; a true
// merge marks v in nextSet, as a true combine() does in the push form.
//   T      accumulator;  id() its identity;  op(a, b) the fold;
//   msg    the message of active neighbour u (w = the edge's weight);
//   fin    merges the folded value into C[v]; any = at least one message;
//   skip   v cannot change this superstep (no need to read its row);
//   kFirst the first active neighbour in row order decides (first writer);
//   kFuseApply  apply(v) may run in the pulling thread (see PullFirstU64);
//   kStale an inactive neighbour's message was already merged into C[v] by the
//          superstep after its last change, so folding it again cannot change
//          fin's result -- true for min-label CC when every vertex began
//          active (labels only fall, C[v] <= every label ever sent to v), and
//          the pull then skips the frontier test.
struct PullSumF64 {                  // PageRank, PageRank_delta, BC Backward: locked +=
  using T = double;
  static constexpr bool kFirst = false;
  static constexpr bool kStale = false;
  static constexpr bool kFuseApply = false;
  __device__ static T id() { return 0.0; }
  __device__ static T msg(const EngCtx& c, uint32_t u, uint32_t) { return ((const double*)c.S)[u]; }
  __device__ static T op(T a, T b) { return a + b; }
  __device__ static bool skip(const EngCtx&, uint32_t) { return false; }
  __device__ static bool fin(const EngCtx& c, uint32_t v, T acc, bool any) {
    if (!any) return false;
    ((double*)c.C)[v] += acc;
    return true;
  }
};
struct PullSumI32 {                  // BC Forward, CountIn: int += (wraps like GRE's int)
  using T = int32_t;
  static constexpr bool kFirst = false;
  static constexpr bool kStale = false;
  static constexpr bool kFuseApply = false;
  __device__ static T id() { return 0; }
  __device__ static T msg(const EngCtx& c, uint32_t u, uint32_t) { return ((const int32_t*)c.S)[u]; }
  __device__ static T op(T a, T b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
  __device__ static bool skip(const EngCtx&, uint32_t) { return false; }
  __device__ static bool fin(const EngCtx& c, uint32_t v, T acc, bool any) {
    if (!any) return false;
    int32_t* cm = (int32_t*)c.C;
    cm[v] = op(cm[v], acc);
    return true;
  }
};
struct PullMinU64 {                  // ConnectedComponents: min label
  using T = unsigned long long;
  static constexpr bool kFirst = false;
  static constexpr bool kStale = true;
  static constexpr bool kFuseApply = false;
  __device__ static T id() { return ~0ULL; }
  __device__ static T msg(const EngCtx& c, uint32_t u, uint32_t) { return ((const unsigned long long*)c.S)[u]; }
  __device__ static T op(T a, T b) { return a < b ? a : b; }
  __device__ static bool skip(const EngCtx&, uint32_t) { return false; }
  __device__ static bool fin(const EngCtx& c, uint32_t v, T acc, bool) {
    unsigned long long* cm = (unsigned long long*)c.C;
    if (!(acc < cm[v])) return false;
    cm[v] = acc;
    return true;
  }
};
struct PullFirstU64 {                // BFS: the first message into an unset combine slot wins
  using T = unsigned long long;
  static constexpr bool kFirst = true;
  static constexpr bool kStale = false;
  // apply(v) runs in the pulling thread right after v's combine, and the pull
  // marks only the vertices it activates: v's combine is complete there (no
  // other thread writes C[v] in the pull form), and BFS's apply writes v's
  // scatter data only when v was never reached, i.e. not active, so no thread
  // reads it as a message in this superstep
  static constexpr bool kFuseApply = true;
  __device__ static T id() { return ~0ULL; }
  __device__ static T msg(const EngCtx& c, uint32_t u, uint32_t) { return ((const unsigned long long*)c.S)[u]; }
  __device__ static T op(T a, T b) { return a != ~0ULL ? a : b; }
  __device__ static bool skip(const EngCtx& c, uint32_t v) { return ((const unsigned long long*)c.C)[v] != ~0ULL; }
  __device__ static bool fin(const EngCtx& c, uint32_t v, T acc, bool any) {
    unsigned long long* cm = (unsigned long long*)c.C;
    if (!any || cm[v] != ~0ULL) return false;
    cm[v] = acc;
    return true;
  }
};
struct PullMinU32W {                 // SSSP: min of lastStates[u] + edgeWgt
  using T = uint32_t;
  static constexpr bool kFirst = false;
  static constexpr bool kStale = false;
  static constexpr bool kFuseApply = false;
  __device__ static T id() { return 0xffffffffu; }
  __device__ static T msg(const EngCtx& c, uint32_t u, uint32_t w) { return ((const uint32_t*)c.S)[u] + w; }
  __device__ static T op(T a, T b) { return a < b ? a : b; }
  __device__ static bool skip(const EngCtx&, uint32_t) { return false; }
  __device__ static bool fin(const EngCtx& c, uint32_t v, T acc, bool any) {
    uint32_t* st = (uint32_t*)c.C;
    if (!any || !(acc < st[v])) return false;
    st[v] = acc;
    return true;
  }
};
template <int P> struct Pull;
template <> struct Pull<GRE_PROG_PAGERANK> : PullSumF64 {};
template <> struct Pull<GRE_PROG_PAGERANK_DELTA> : PullSumF64 {};
template <> struct Pull<GRE_PROG_BC_BACKWARD> : PullSumF64 {};
template <> struct Pull<GRE_PROG_BC_FORWARD> : PullSumI32 {};
template <> struct Pull<GRE_PROG_COUNTIN> : PullSumI32 {};
template <> struct Pull<GRE_PROG_CC> : PullMinU64 {};
template <> struct Pull<GRE_PROG_BFS> : PullFirstU64 {};
template <> struct Pull<GRE_PROG_SSSP> : PullMinU32W {};

// v's pull row: for Out, its in-edges (CSC row, an active u's out-edge u -> v);
// for In, its out-edges (CSR row, an active u's in-edge v -> u); InOut: CSC then CSR
struct PullAdj {
  const uint64_t* off; const uint32_t* idx;     // CSR
  const uint64_t* coff; const uint32_t* cidx;   // CSC
  const uint32_t* w_csr; const uint32_t* w_csc; // SSSP weights in CSR / CSC order, or nullptr
  int dir;
};

__device__ inline uint64_t pull_deg(const PullAdj& a, uint32_t v) {
  uint64_t d = 0;
  if (a.dir != GRE_DIR_IN) d += a.coff[v + 1] - a.coff[v];
  if (a.dir != GRE_DIR_OUT) d += a.off[v + 1] - a.off[v];
  return d;
}

__device__ inline void pull_entry(const PullAdj& a, uint32_t v, uint64_t j, uint32_t& u, uint32_t& w) {
  const uint64_t din = a.dir != GRE_DIR_IN ? a.coff[v + 1] - a.coff[v] : 0;
  if (j < din) {
    const uint64_t p = a.coff[v] + j;
    u = a.cidx[p];
    w = a.w_csc ? a.w_csc[p] : 0u;
  } else {
    const uint64_t p = a.off[v] + (j - din);
    u = a.idx[p];
    w = a.w_csr ? a.w_csr[p] : 0u;
  }
}

constexpr uint64_t kPullShort = 32;      // rows up to this many entries: one thread
constexpr uint64_t kPullShortFirst = 1024;  // the same for first-writer programs (most rows stop early)
constexpr uint64_t kPullLong = 65536;    // rows above this: one 1024-thread block; between: one wave
template <typename R>
constexpr uint64_t short_limit() { return R::kFirst ? kPullShortFirst : kPullShort; }

__device__ inline bool bit_test(const uint32_t* b, uint32_t v) { return (b[v >> 5] >> (v & 31)) & 1u; }

__device__ inline void mark_next(uint32_t* next, uint32_t v) {
  const uint32_t bit = 1u << (v & 31);
  if (!(next[v >> 5] & bit)) atomicOr(next + (v >> 5), bit);
}

// folds the active entries [b, e) of one adjacency array into acc, in entry
// order; false once a first-writer program has its message.  Eight entries per
// step: their ids, then their frontier bits (and, for folding programs, their
// messages) are loaded together, so a step waits for two or three memory
// latencies instead of two per entry.  Past the row's end a step reads entry b
// again (a valid address) and ignores it.
template <typename R>
__device__ inline bool fold_run(const EngCtx& c, const uint32_t* __restrict__ idx, const uint32_t* __restrict__ w,
                                uint64_t b, uint64_t e, const uint32_t* __restrict__ curr, bool all,
                                typename R::T& acc, bool& any) {
  using T = typename R::T;
  for (uint64_t p = b; p < e; p += 8) {
    uint64_t q[8];
    uint32_t u[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) q[k] = p + k < e ? p + k : b;
#pragma unroll
    for (int k = 0; k < 8; ++k) u[k] = idx[q[k]];
    bool on[8];
    if (all) {
#pragma unroll
      for (int k = 0; k < 8; ++k) on[k] = p + k < e;
    } else {
      uint32_t fw[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) fw[k] = curr[u[k] >> 5];
#pragma unroll
      for (int k = 0; k < 8; ++k) on[k] = p + k < e && ((fw[k] >> (u[k] & 31)) & 1u);
    }
    if (R::kFirst) {
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (on[k]) {
          acc = R::op(acc, R::msg(c, u[k], w ? w[q[k]] : 0u));
          any = true;
          return false;
        }
    } else {
      T m[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) m[k] = R::msg(c, u[k], w ? w[q[k]] : 0u);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (on[k]) {
          acc = R::op(acc, m[k]);
          any = true;
        }
    }
  }
  return true;
}

template <int P>
__global__ __launch_bounds__(kB) void k_pull_short(EngCtx c, PullAdj a, uint32_t nV,
                                                   const uint32_t* __restrict__ curr, bool all,
                                                   uint32_t* __restrict__ next) {
  using R = Pull<P>;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV; v += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t b0 = 0, e0 = 0, b1 = 0, e1 = 0;
    if (a.dir != GRE_DIR_IN) { b0 = a.coff[v]; e0 = a.coff[v + 1]; }
    if (a.dir != GRE_DIR_OUT) { b1 = a.off[v]; e1 = a.off[v + 1]; }
    const uint64_t d = (e0 - b0) + (e1 - b1);
    if (d > short_limit<R>() || d == 0 || R::skip(c, (uint32_t)v)) continue;
    typename R::T acc = R::id();
    bool any = false;
    // CSC part first, then CSR (the row order of pull_entry)
    if (fold_run<R>(c, a.cidx, a.w_csc, b0, e0, curr, all, acc, any))
      fold_run<R>(c, a.idx, a.w_csr, b1, e1, curr, all, acc, any);
    if (R::fin(c, (uint32_t)v, acc, any) && (!R::kFuseApply || Prog<P>::apply(c, (uint32_t)v)))
      mark_next(next, (uint32_t)v);
  }
}

// one wave (or one block of W waves) per listed row; the fold runs over lanes
// in a fixed order, and for first-writer programs the smallest row position
// with an active neighbour decides
template <typename R>
__device__ inline void wave_fold(typename R::T& acc, uint64_t& first) {
  for (int o = 32; o > 0; o >>= 1) {
    const typename R::T oa = __shfl_xor(acc, o, 64);
    const uint64_t of = __shfl_xor(first, o, 64);
    if (R::kFirst) {
      if (of < first) { first = of; acc = oa; }
    } else {
      acc = R::op(acc, oa);
      first = of < first ? of : first;
    }
  }
}

// W == 1: blocks of kB threads, one wave per row; W > 1: blocks of 64*W threads, one block per row
template <int P, int W>
__global__ __launch_bounds__(W == 1 ? kB : 64 * W) void k_pull_rows(EngCtx c, PullAdj a, const uint32_t* __restrict__ rows,
                                                      uint32_t nrows, const uint32_t* __restrict__ curr,
                                                      bool all, uint32_t* __restrict__ next) {
  using R = Pull<P>;
  using T = typename R::T;
  __shared__ T sacc[W];
  __shared__ uint64_t sfirst[W];
  const int lane = (int)(threadIdx.x & 63), wv = (int)(threadIdx.x >> 6);
  const uint64_t per = W == 1 ? (uint64_t)(blockDim.x >> 6) : 1;
  const uint64_t r0 = W == 1 ? (uint64_t)blockIdx.x * per + wv : blockIdx.x;
  const uint64_t rstep = W == 1 ? (uint64_t)gridDim.x * per : gridDim.x;
  for (uint64_t r = r0; r < nrows; r += rstep) {
    const uint32_t v = rows[r];
    const bool skip = R::skip(c, v);
    const uint64_t d = skip ? 0 : pull_deg(a, v);
    T acc = R::id();
    uint64_t first = ~0ULL;
    const uint64_t t0 = W == 1 ? (uint64_t)lane : threadIdx.x;
    const uint64_t ts = W == 1 ? 64 : 64 * W;
    for (uint64_t base = 0; base < d; base += ts) {
      const uint64_t j = base + t0;
      if (j < d && (!R::kFirst || first == ~0ULL)) {
        uint32_t u, w;
        pull_entry(a, v, j, u, w);
        if (all || bit_test(curr, u)) {
          if (R::kFirst) { if (first == ~0ULL) { first = j; acc = R::msg(c, u, w); } }
          else { acc = R::op(acc, R::msg(c, u, w)); first = j < first ? j : first; }
        }
      }
      if (R::kFirst) {
        // stop at the first chunk that found one (every lane sees the same chunks)
        bool found = __any(first != ~0ULL);
        if (W > 1) {
          __syncthreads();
          if (lane == 0) sfirst[wv] = found ? 1ULL : 0ULL;
          __syncthreads();
          found = false;
          for (int q = 0; q < W; ++q) found = found || sfirst[q] != 0ULL;
          __syncthreads();
        }
        if (found) break;
      }
    }
    wave_fold<R>(acc, first);
    if (W > 1) {
      if (lane == 0) { sacc[wv] = acc; sfirst[wv] = first; }
      __syncthreads();
      if (threadIdx.x == 0) {
        for (int q = 1; q < W; ++q) {
          if (R::kFirst) {
            if (sfirst[q] < first) { first = sfirst[q]; acc = sacc[q]; }
          } else {
            acc = R::op(acc, sacc[q]);
            first = sfirst[q] < first ? sfirst[q] : first;
          }
        }
      }
      __syncthreads();
    }
    if ((W == 1 ? lane : (int)threadIdx.x) == 0 && !skip) {
      if (R::fin(c, v, acc, first != ~0ULL) && (!R::kFuseApply || Prog<P>::apply(c, v))) mark_next(next, v);
    }
  }
}

// rows of pull degree in (lim, kPullLong] and above kPullLong
__global__ void k_pull_classify(PullAdj a, uint32_t nV, uint64_t lim, uint32_t* __restrict__ mid,
                                uint32_t* __restrict__ nmid, uint32_t* __restrict__ lng, uint32_t* __restrict__ nlng) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t nvpad = (nV + 63) & ~63ULL;   // whole waves stay in the loop for the ballots
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nvpad; v += stride) {
    const uint64_t d = v < nV ? pull_deg(a, (uint32_t)v) : 0;
    wave_append(d > kPullLong, (uint32_t)v, lng, nlng);
    wave_append(d > lim && d <= kPullLong, (uint32_t)v, mid, nmid);
  }
}

// Complete frontier (every vertex with edges in the push direction is active)
// and an FP64 sum combine: the pull is the tile-segmented reduction over the
// pull adjacency (one array: CSC for Out, CSR for In), summed in a fixed order;
// every vertex with a non-empty pull row receives a message.
struct EngSumMode {
  using T = double;
  const double* S;
  double* C;
  __device__ double gather(uint64_t, uint32_t u) const { return S[u]; }
  __device__ double pre(uint32_t r) const { return C[r]; }
  __device__ void finalize_pre(uint32_t r, double s, double c, double&, double&) const { C[r] = c + s; }
  __device__ void finalize(uint32_t r, double s, double& a0, double& a1) const { finalize_pre(r, s, pre(r), a0, a1); }
};

// The same over the XCD source partition of the CSC (unweighted PageRank's
// partitioned pull): the scatter data is permuted to hub-sorted compact source
// order, each (row, source range) segment is summed, then each row adds its
// segments' sums, all in a fixed order.
struct EngSegMode {
  using T = double;
  const double* gin;   // scatter data in compact source order
  double* sum;         // per segment
  __device__ double gather(uint64_t, uint32_t u) const { return gin[u]; }
  __device__ void finalize(uint32_t seg, double s, double&, double&) const { sum[seg] = s; }
};
struct EngRowMode {
  using T = double;
  const double* seg;   // segment sums
  double* C;
  __device__ double gather(uint64_t, uint32_t sg) const { return seg[sg]; }
  __device__ double pre(uint32_t r) const { return C[r]; }
  __device__ void finalize_pre(uint32_t r, double s, double c, double&, double&) const { C[r] = c + s; }
  __device__ void finalize(uint32_t r, double s, double& a0, double& a1) const { finalize_pre(r, s, pre(r), a0, a1); }
};

// cid[v] is v's compact id if v has out-edges, else 0xffffffff (no entry)
__global__ void k_to_compact(const double* __restrict__ S, const uint32_t* __restrict__ cid, uint32_t nV,
                             double* __restrict__ Sc) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nV; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t k = cid[v];
    if (k < nV) Sc[k] = S[v];
  }
}

// nextSet = vertices with a non-empty row in off (whole words, as k_init_all)
__global__ void k_mark_rows(const uint64_t* __restrict__ off, uint32_t nV, uint32_t nvpad,
                            uint32_t* __restrict__ next) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nvpad;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const bool on = v < nV && off[v + 1] > off[v];
    const uint64_t m = __ballot(on);
    if ((threadIdx.x & 63) == 0) {
      next[v >> 5] = (uint32_t)m;
      next[(v >> 5) + 1] = (uint32_t)(m >> 32);
    }
  }
}

// ---- frontier plumbing ------------------------------------------------------
struct Adj {
  const uint64_t* off; const uint32_t* idx;     // CSR
  const uint64_t* coff; const uint32_t* cidx;   // CSC
  const uint64_t* ceid;                         // CSC -> CSR position
  int dir;
};

__device__ inline uint64_t degree(const Adj& a, uint32_t v) {
  uint64_t d = 0;
  if (a.dir != GRE_DIR_IN) d += a.off[v + 1] - a.off[v];
  if (a.dir != GRE_DIR_OUT) d += a.coff[v + 1] - a.coff[v];
  return d;
}

template <int P>
__global__ void k_init_all(EngCtx c, uint32_t nV, uint32_t nvpad, uint32_t* __restrict__ curr) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < nvpad;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const bool on = v < nV && Prog<P>::init(c, (uint32_t)v);
    const uint64_t m = __ballot(on);
    if ((threadIdx.x & 63) == 0) {
      curr[v >> 5] = (uint32_t)m;
      curr[(v >> 5) + 1] = (uint32_t)(m >> 32);
    }
  }
}

// master pre-insert of the sources, then init() on each (Engine.hpp:252-257, :459-471)
template <int P>
__global__ void k_init_src(EngCtx c, const uint32_t* __restrict__ src, int64_t n, uint32_t* __restrict__ curr) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int64_t i = 0; i < n; ++i) curr[src[i] >> 5] |= 1u << (src[i] & 31);
  for (int64_t i = 0; i < n; ++i)
    if (Prog<P>::init(c, src[i])) curr[src[i] >> 5] |= 1u << (src[i] & 31);
}

// ---- active-set kernels ---------------------------------------------------
// One 256-thread block per kVB vertices (kVW bitmap words).  Work is spread
// over the set bits only: the block takes the popcount prefix of its words,
// then thread j handles the j-th set bit (word by binary search on the prefix,
// bit by select), so a sparse bitmap costs little more than reading it and a
// dense one keeps every thread busy.
constexpr int kVW = kVB / 32;

// the r-th (0-based) set bit of m; r < popc(m)
__device__ inline uint32_t select_bit(uint32_t m, uint32_t r) {
  uint32_t pos = 0, c;
  c = (uint32_t)__popc(m & 0xffffu); if (r >= c) { r -= c; pos += 16; m >>= 16; }
  c = (uint32_t)__popc(m & 0xffu);   if (r >= c) { r -= c; pos += 8;  m >>= 8; }
  c = (uint32_t)__popc(m & 0xfu);    if (r >= c) { r -= c; pos += 4;  m >>= 4; }
  c = (uint32_t)__popc(m & 0x3u);    if (r >= c) { r -= c; pos += 2;  m >>= 2; }
  if (r >= (m & 1u)) pos += 1;
  return pos;
}

// words[0..kVW) in LDS -> pre[0..kVW] exclusive popcount prefix (pre[kVW] = total);
// every thread of the block calls it
__device__ inline void word_prefix(const uint32_t* words, uint32_t* pre, uint32_t* wtot) {
  const int t = (int)threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t x = t < kVW ? (uint32_t)__popc(words[t]) : 0u;
  const uint32_t own = x;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63 && w < kVW / 64) wtot[w] = x;
  __syncthreads();
  if (t < kVW) {
    uint32_t add = 0;
    for (int k = 0; k < w; ++k) add += wtot[k];
    pre[t] = add + x - own;
    if (t == kVW - 1) pre[kVW] = add + x;
  }
  __syncthreads();
}

// word index of the j-th set bit: the last wi with pre[wi] <= j
__device__ inline int word_of(const uint32_t* pre, uint32_t j) {
  int lo = 0, hi = kVW;   // pre[lo] <= j < pre[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] <= j) lo = mid; else hi = mid;
  }
  return lo;
}

// number of active vertices of the block and their degree sum, from the
// block's active words in LDS (pre = their prefix), written to bcnt/bdeg
__device__ inline void block_count_active(const uint32_t* act, const uint32_t* pre, uint32_t w0, Adj a,
                                          uint32_t* __restrict__ bcnt, uint64_t* __restrict__ bdeg, uint64_t* red) {
  const uint32_t total = pre[kVW];
  uint64_t deg = 0;
  for (uint32_t j = threadIdx.x; j < total; j += kB) {
    const int wi = word_of(pre, j);
    const uint32_t v = (w0 + (uint32_t)wi) * 32u + select_bit(act[wi], j - pre[wi]);
    deg += degree(a, v);
  }
  for (int o = 32; o > 0; o >>= 1) deg += __shfl_xor(deg, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = deg;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t d = 0;
    for (int w = 0; w < kB / 64; ++w) d += red[w];
    bcnt[blockIdx.x] = total;
    bdeg[blockIdx.x] = d;
  }
}

// count and degree sum of the active vertices (curr) per block
__global__ __launch_bounds__(kB) void k_count(const uint32_t* __restrict__ curr, uint32_t nV, Adj a,
                                              uint32_t* __restrict__ bcnt, uint64_t* __restrict__ bdeg) {
  __shared__ uint32_t act[kVW];
  __shared__ uint32_t pre[kVW + 1];
  __shared__ uint32_t wtot[kVW / 64];
  __shared__ uint64_t red[kB / 64];
  const uint32_t w0 = blockIdx.x * (uint32_t)kVW;
  (void)nV;   // bits past nV are never set
  if (threadIdx.x < kVW) act[threadIdx.x] = curr[w0 + threadIdx.x];
  __syncthreads();
  word_prefix(act, pre, wtot);
  block_count_active(act, pre, w0, a, bcnt, bdeg, red);
}

// apply + halt + count: new curr bit = (kHalt ? 0 : curr bit) | (next bit && apply(v)).
// apply(v) touches only v's own data, so the order of the applies is free.
template <int P>
__global__ __launch_bounds__(kB) void k_apply(EngCtx c, const uint32_t* __restrict__ next, uint32_t* __restrict__ curr,
                                              uint32_t nV, Adj a, uint32_t* __restrict__ bcnt,
                                              uint64_t* __restrict__ bdeg) {
  __shared__ uint32_t nxt[kVW];
  __shared__ uint32_t act[kVW];
  __shared__ uint32_t pre[kVW + 1];
  __shared__ uint32_t wtot[kVW / 64];
  __shared__ uint64_t red[kB / 64];
  const uint32_t w0 = blockIdx.x * (uint32_t)kVW;
  (void)nV;   // next holds no bit past nV
  if (threadIdx.x < kVW) {
    nxt[threadIdx.x] = next[w0 + threadIdx.x];
    act[threadIdx.x] = Prog<P>::kHalt ? 0u : curr[w0 + threadIdx.x];
  }
  __syncthreads();
  word_prefix(nxt, pre, wtot);
  const uint32_t total = pre[kVW];
  for (uint32_t j = threadIdx.x; j < total; j += kB) {
    const int wi = word_of(pre, j);
    const uint32_t b = select_bit(nxt[wi], j - pre[wi]);
    if (Prog<P>::apply(c, (w0 + (uint32_t)wi) * 32u + b)) atomicOr(&act[wi], 1u << b);
  }
  __syncthreads();
  if (threadIdx.x < kVW) curr[w0 + threadIdx.x] = act[threadIdx.x];
  word_prefix(act, pre, wtot);
  block_count_active(act, pre, w0, a, bcnt, bdeg, red);
}

// after a pull that applied in place (kFuseApply): next holds exactly the
// vertices apply() activated; new curr = (kHalt ? 0 : curr) | next, counted
template <int P>
__global__ __launch_bounds__(kB) void k_activate(const uint32_t* __restrict__ next, uint32_t* __restrict__ curr, Adj a,
                                                 uint32_t* __restrict__ bcnt, uint64_t* __restrict__ bdeg) {
  __shared__ uint32_t act[kVW];
  __shared__ uint32_t pre[kVW + 1];
  __shared__ uint32_t wtot[kVW / 64];
  __shared__ uint64_t red[kB / 64];
  const uint32_t w0 = blockIdx.x * (uint32_t)kVW;
  if (threadIdx.x < kVW) {
    const uint32_t m = next[w0 + threadIdx.x] | (Prog<P>::kHalt ? 0u : curr[w0 + threadIdx.x]);
    act[threadIdx.x] = m;
    curr[w0 + threadIdx.x] = m;
  }
  __syncthreads();
  word_prefix(act, pre, wtot);
  block_count_active(act, pre, w0, a, bcnt, bdeg, red);
}

// exclusive scan of block (count, degree) pairs; totals at [nb] and in
// tot[0..1] (one read-back).  One 1024-thread block, kScanPer consecutive
// pairs per thread and pass: every load of a pass is issued before any is
// used, the per-thread sums are scanned with wave shuffles and one LDS step.
constexpr int kScanPer = 16;
__global__ __launch_bounds__(1024) void k_scan_blocks(uint32_t* __restrict__ bcnt, uint64_t* __restrict__ bdeg,
                                                      uint32_t nb, uint64_t* __restrict__ tot) {
  __shared__ uint32_t wc[16];
  __shared__ uint64_t wd[16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t carry_c = 0;
  uint64_t carry_d = 0;
  for (uint32_t base = 0; base < nb; base += 1024 * kScanPer) {
    const uint32_t b = base + threadIdx.x * kScanPer;
    uint32_t cv[kScanPer];
    uint64_t dv[kScanPer];
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
      const uint32_t i = b + k < nb ? b + k : nb - 1;   // a valid pair, ignored below
      cv[k] = bcnt[i];
      dv[k] = bdeg[i];
    }
    uint32_t c = 0;
    uint64_t d = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k)
      if (b + k < nb) { c += cv[k]; d += dv[k]; }
    uint32_t ic = c;
    uint64_t id = d;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t tc = __shfl_up(ic, o, 64);
      const uint64_t td = __shfl_up(id, o, 64);
      if (lane >= o) { ic += tc; id += td; }
    }
    if (lane == 63) { wc[w] = ic; wd[w] = id; }
    __syncthreads();
    uint32_t rc = carry_c + ic - c;
    uint64_t rd = carry_d + id - d;
    for (int k = 0; k < w; ++k) { rc += wc[k]; rd += wd[k]; }
#pragma unroll
    for (int k = 0; k < kScanPer; ++k)
      if (b + k < nb) {
        bcnt[b + k] = rc;
        bdeg[b + k] = rd;
        rc += cv[k];
        rd += dv[k];
      }
    for (int k = 0; k < 16; ++k) { carry_c += wc[k]; carry_d += wd[k]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    bcnt[nb] = carry_c;
    bdeg[nb] = carry_d;
    tot[0] = carry_c;
    tot[1] = carry_d;
  }
}

// queue[pos] = v, qpos[pos] = edge prefix; qpos[total] = total edges
__global__ __launch_bounds__(kB) void k_write(const uint32_t* __restrict__ curr, uint32_t nV, Adj a,
                                               const uint32_t* __restrict__ bcnt, const uint64_t* __restrict__ bdeg,
                                               uint32_t nb, uint32_t* __restrict__ queue,
                                               uint64_t* __restrict__ qpos) {
  __shared__ uint32_t wc[kB / 64];
  __shared__ uint64_t wd[kB / 64];
  const uint64_t v0 = (uint64_t)blockIdx.x * kVB;
  uint32_t base_c = bcnt[blockIdx.x];
  uint64_t base_d = bdeg[blockIdx.x];
  if (blockIdx.x == 0 && threadIdx.x == 0) qpos[bcnt[nb]] = bdeg[nb];
  // bcnt holds the exclusive prefix of the blocks' counts: a block with no
  // active vertex has nothing to write (small frontiers skip most blocks)
  if (bcnt[blockIdx.x + 1] == base_c) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i0 = 0; i0 < kVB; i0 += kB) {
    const uint64_t v = v0 + i0 + threadIdx.x;
    const bool on = v < nV && bit_test(curr, (uint32_t)v);
    const uint64_t d = on ? degree(a, (uint32_t)v) : 0;
    // wave-inclusive scans of (on, d)
    uint32_t c = on ? 1u : 0u;
    uint64_t dd = d;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t tc = __shfl_up(c, o, 64);
      const uint64_t td = __shfl_up(dd, o, 64);
      if (lane >= o) { c += tc; dd += td; }
    }
    if (lane == 63) { wc[w] = c; wd[w] = dd; }
    __syncthreads();
    uint32_t pc = base_c;
    uint64_t pd = base_d;
    for (int q = 0; q < w; ++q) { pc += wc[q]; pd += wd[q]; }
    if (on) {
      const uint32_t pos = pc + c - 1;
      queue[pos] = (uint32_t)v;
      qpos[pos] = pd + dd - d;
    }
    uint32_t tc = 0; uint64_t td = 0;
    for (int q = 0; q < kB / 64; ++q) { tc += wc[q]; td += wd[q]; }
    base_c += tc;
    base_d += td;
    __syncthreads();
  }
}

// edge-balanced push: virtual edge k of the frontier -> (u, v, CSR position)
template <int P>
__global__ __launch_bounds__(kB) void k_expand(EngCtx c, Adj a, const uint32_t* __restrict__ queue,
                                               const uint64_t* __restrict__ qpos, uint32_t qn, uint64_t T,
                                               uint32_t* __restrict__ next) {
  __shared__ uint32_t range[2];
  const uint64_t k0 = (uint64_t)blockIdx.x * kEB;
  const uint64_t k1 = k0 + kEB < T ? k0 + kEB : T;
  if (threadIdx.x < 2) {
    const uint64_t key = threadIdx.x == 0 ? k0 : k1 - 1;
    uint32_t lo = 0, hi = qn;  // last i with qpos[i] <= key
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (qpos[mid] <= key) lo = mid; else hi = mid;
    }
    range[threadIdx.x] = lo;
  }
  __syncthreads();
  const uint32_t i0 = range[0], i1 = range[1];
  for (uint64_t k = k0 + threadIdx.x; k < k1; k += kB) {
    uint32_t lo = i0, hi = i1 + 1;
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (qpos[mid] <= k) lo = mid; else hi = mid;
    }
    const uint32_t u = queue[lo];
    uint64_t j = k - qpos[lo];
    uint32_t v;
    uint64_t epos;
    const uint64_t dout = a.dir == GRE_DIR_IN ? 0 : a.off[u + 1] - a.off[u];
    if (j < dout) {
      epos = a.off[u] + j;
      v = a.idx[epos];
    } else {
      const uint64_t p = a.coff[u] + (j - dout);
      v = a.cidx[p];
      epos = a.ceid ? a.ceid[p] : p;
    }
    if (Prog<P>::combine(c, u, v, epos)) {
      const uint32_t bit = 1u << (v & 31);
      if (!(next[v >> 5] & bit)) atomicOr(next + (v >> 5), bit);
    }
  }
}

__global__ void k_zero_u32(uint32_t* p, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

template <typename T>
__global__ void k_fill(T* p, uint64_t n, T v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

unsigned gcap(uint64_t n) {
  uint64_t g = (n + kB - 1) / kB;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

}  // namespace
}  // namespace gre

struct gre_engine {
  gre_graph* g = nullptr;
  int prog = 0;
  int64_t max_ss = 1024;
  gre::DevBuf S, C, X, AUX;
  size_t es[4] = {0, 0, 0, 0};
  bool c_is_x = false;  // SSSP: combine and vertex data are one array
  gre::DevBuf curr, next, queue, qpos, bcnt, bdeg, tot;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};   // compute start, compute end, apply end
  hipEvent_t ev_tot = nullptr;                       // the active totals have reached the host
  uint64_t* tot_host = nullptr;                      // pinned: (active count, active edges)
  uint32_t nwords = 0, nb = 0;
  int32_t level = 1;
  int dir = GRE_DIR_OUT;
  bool begun = false;
  int64_t superstep = 0;
  uint32_t qn = 0;
  uint64_t qedges = 0;
  bool queue_stale = false;    // counts and prefixes are current, the queue itself is not written yet
  bool sc_valid = false;       // Sc holds the current scatter data (GRE's PageRank never changes it after init)
  gre::DevBuf rows_mark;       // nextSet of the rows with a non-empty CSC row (the partitioned pull's)
  double times[5] = {0, 0, 0, 0, 0};
  int mode = 0;                // GRE_ENGINE_AUTO / _PUSH / _PULL
  int64_t pull_steps = 0;      // supersteps of the last run done in the pull form
  uint64_t seen_edges = 0;     // first-writer programs: edges of every frontier so far
  bool pulling = false;        // first-writer programs: the last superstep pulled
  bool all_began = false;      // the run began with every vertex active (no sources)
  int sum_dir = -1;            // direction of the tile plan below
  gre::PullPlan sum_plan;      // tile-segmented pull of complete-frontier FP64 sums
  gre::PullBuffers sum_buf;
  // the same over the XCD source partition (Out, large graphs)
  std::shared_ptr<gre::XPart> xp;
  gre::PullPlan seg_plan, row_plan;
  gre::PullBuffers seg_buf, row_buf;
  gre::DevBuf Sc, xsum;
  // the binned FP64 sum pull (GRE's PageRank, complete frontier, large graphs)
  std::shared_ptr<gre::XPart> bxp;
  gre::EngBinBuffers bbuf;
  bool bsc_valid = false;      // bbuf holds the current scatter data's compact copy
};

namespace gre {
namespace {

EngCtx make_ctx(gre_engine* e) {
  EngCtx c;
  c.S = e->S.p;
  c.C = e->c_is_x ? e->X.p : e->C.p;
  c.X = e->X.p;
  c.aux = e->AUX.p;
  c.level = e->level;
  c.outdeg = e->g->outdeg.as<int32_t>();
  c.w_csr = e->g->w_u32.as<uint32_t>();
  return c;
}

Adj make_adj(gre_engine* e) {
  Adj a;
  gre_graph* g = e->g;
  a.off = g->off.as<uint64_t>();
  a.idx = g->idx.as<uint32_t>();
  a.coff = g->coff.as<uint64_t>();
  a.cidx = g->cidx.as<uint32_t>();
  a.ceid = (e->prog == GRE_PROG_SSSP && e->dir != GRE_DIR_OUT) ? graph_ceid(g) : nullptr;
  a.dir = e->dir;
  return a;
}

PullAdj make_pull_adj(gre_engine* e) {
  PullAdj a;
  gre_graph* g = e->g;
  a.off = g->off.as<uint64_t>();
  a.idx = g->idx.as<uint32_t>();
  a.coff = g->coff.as<uint64_t>();
  a.cidx = g->cidx.as<uint32_t>();
  const bool w = e->prog == GRE_PROG_SSSP;
  a.w_csr = (w && e->dir != GRE_DIR_OUT) ? g->w_u32.as<uint32_t>() : nullptr;
  a.w_csc = (w && e->dir != GRE_DIR_IN) ? graph_w_csc(g) : nullptr;
  a.dir = e->dir;
  return a;
}

// rows too long for one thread, listed once per graph, direction and short-row
// limit (first-writer programs keep longer rows on one thread)
const PullRows& ensure_pull_rows(gre_engine* e, const PullAdj& a, uint64_t lim, hipStream_t s) {
  PullRows& r = e->g->pull_rows[e->dir + (lim == kPullShort ? 0 : 4)];
  if (r.mid.p) return r;
  const uint32_t nV = e->g->nV;
  DevBuf cnt;
  cnt.alloc(8);
  r.mid.alloc(4 * (size_t)nV);
  r.lng.alloc(4 * (size_t)nV);
  GRE_HIP(hipMemsetAsync(cnt.p, 0, 8, s));
  hipLaunchKernelGGL(k_pull_classify, dim3(gcap(nV)), dim3(kB), 0, s, a, nV, lim, r.mid.as<uint32_t>(),
                     cnt.as<uint32_t>(), r.lng.as<uint32_t>(), cnt.as<uint32_t>() + 1);
  GRE_HIP(hipGetLastError());
  uint32_t h[2];
  GRE_HIP(hipMemcpyAsync(h, cnt.p, 8, hipMemcpyDeviceToHost, s));
  GRE_HIP(hipStreamSynchronize(s));
  r.n_mid = h[0];
  r.n_long = h[1];
  return r;
}

template <typename F>
void dispatch(int prog, F&& f) {
  switch (prog) {
    case GRE_PROG_PAGERANK: f(std::integral_constant<int, GRE_PROG_PAGERANK>()); break;
    case GRE_PROG_PAGERANK_DELTA: f(std::integral_constant<int, GRE_PROG_PAGERANK_DELTA>()); break;
    case GRE_PROG_BFS: f(std::integral_constant<int, GRE_PROG_BFS>()); break;
    case GRE_PROG_CC: f(std::integral_constant<int, GRE_PROG_CC>()); break;
    case GRE_PROG_SSSP: f(std::integral_constant<int, GRE_PROG_SSSP>()); break;
    case GRE_PROG_BC_FORWARD: f(std::integral_constant<int, GRE_PROG_BC_FORWARD>()); break;
    case GRE_PROG_BC_BACKWARD: f(std::integral_constant<int, GRE_PROG_BC_BACKWARD>()); break;
    case GRE_PROG_COUNTIN: f(std::integral_constant<int, GRE_PROG_COUNTIN>()); break;
    default: throw Error(GRE_DEV_NULL, "unknown vertex program id");
  }
}

// Fill the arrays exactly as the program's alloc() leaves them
// (include/Kernels/*.hpp alloc(), VertexCode.hpp:132-178).
void default_state(gre_engine* e) {
  hipStream_t s = stream();
  const uint64_t n = e->g->nV;
  auto fill_u64 = [&](DevBuf& b, uint64_t v) {
    hipLaunchKernelGGL(k_fill<uint64_t>, dim3(gcap(n)), dim3(kB), 0, s, b.as<uint64_t>(), n, v);
  };
  auto zero = [&](DevBuf& b) { GRE_HIP(hipMemsetAsync(b.p, 0, b.bytes, s)); };
  switch (e->prog) {
    case GRE_PROG_PAGERANK: case GRE_PROG_PAGERANK_DELTA:
    case GRE_PROG_BC_BACKWARD: case GRE_PROG_COUNTIN:
      zero(e->S); zero(e->C); zero(e->X); break;
    case GRE_PROG_BFS: case GRE_PROG_CC:
      zero(e->S); fill_u64(e->C, ~0ULL); fill_u64(e->X, ~0ULL); break;
    case GRE_PROG_SSSP:
      zero(e->S);
      hipLaunchKernelGGL(k_fill<uint32_t>, dim3(gcap(n)), dim3(kB), 0, s, e->X.as<uint32_t>(), n, 0xffffffffu);
      break;
    case GRE_PROG_BC_FORWARD: {
      zero(e->S); zero(e->C);
      // vtxData(): level = -1, numPaths = 0
      hipLaunchKernelGGL(k_fill<uint64_t>, dim3(gcap(n)), dim3(kB), 0, s, e->X.as<uint64_t>(), n,
                         (uint64_t)0xffffffffULL);
      break;
    }
  }
  GRE_HIP(hipGetLastError());
}

// The frontier's size and out-degree total, from the blocks' counts.  The
// queue itself (k_write) is written only when a push step needs it: pull
// steps read the currSet bitmap, which stays unchanged until the next apply.
// done (optional) is recorded once the active set is rebuilt, before the
// read-back the host waits for
void rebuild_queue(gre_engine* e, const Adj& a, hipStream_t s, hipEvent_t done = nullptr) {
  (void)a;
  if (!e->tot.p) e->tot.alloc(16);
  hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, e->bcnt.as<uint32_t>(), e->bdeg.as<uint64_t>(), e->nb,
                     e->tot.as<uint64_t>());
  GRE_HIP(hipGetLastError());
  if (done) GRE_HIP(hipEventRecord(done, s));
  e->queue_stale = true;
  // the totals go to pinned host memory and the host polls for them: a
  // superstep waits here every time, and a sleeping wait adds tens of
  // microseconds of wake-up to each
  if (!e->tot_host) GRE_HIP(hipHostMalloc((void**)&e->tot_host, 16, hipHostMallocDefault));
  if (!e->ev_tot) GRE_HIP(hipEventCreateWithFlags(&e->ev_tot, hipEventDisableTiming));
  GRE_HIP(hipMemcpyAsync(e->tot_host, e->tot.p, 16, hipMemcpyDeviceToHost, s));
  GRE_HIP(hipEventRecord(e->ev_tot, s));
  hipError_t q;
  while ((q = hipEventQuery(e->ev_tot)) == hipErrorNotReady) {
  }
  GRE_HIP(q);
  e->qn = (uint32_t)e->tot_host[0];
  e->qedges = e->tot_host[1];
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void engine_begin(gre_engine* e, int dir, const uint32_t* srcs, int64_t nsrc) {
  GRE_CHECK(dir >= GRE_DIR_OUT && dir <= GRE_DIR_INOUT, GRE_DEV_NULL, "dir must be Out, In or InOut");
  GRE_CHECK(nsrc >= 0, GRE_DEV_NULL, "nsrc must be >= 0");
  GRE_CHECK(e->prog != GRE_PROG_SSSP || e->g->w_u32.p, GRE_DEV_NULL,
            "SSSP needs edge weights (gre_graph_set_weights_u32)");
  GRE_CHECK(e->prog != GRE_PROG_BC_BACKWARD || e->AUX.p, GRE_DEV_NULL,
            "BC backward needs the forward state (GRE_ARR_AUX)");
  if (srcs)
    for (int64_t i = 0; i < nsrc; ++i)
      GRE_CHECK(srcs[i] < e->g->nV, GRE_DEV_NULL, "source id >= nV");
  hipStream_t s = stream();
  e->dir = dir;
  e->sc_valid = false;  // init writes the scatter data
  e->bsc_valid = false;
  const double t0 = now();
  // run(): superstep != 0 -> currSet.clear(), nextSet.clear() (Engine.hpp:182-193)
  GRE_HIP(hipMemsetAsync(e->curr.p, 0, e->curr.bytes, s));
  GRE_HIP(hipMemsetAsync(e->next.p, 0, e->next.bytes, s));
  EngCtx c = make_ctx(e);
  const uint32_t nvpad = e->nwords * 32;
  if (!srcs) {
    dispatch(e->prog, [&](auto P) {
      hipLaunchKernelGGL(k_init_all<decltype(P)::value>, dim3(gcap(nvpad)), dim3(kB), 0, s, c, e->g->nV, nvpad,
                         e->curr.as<uint32_t>());
    });
  } else if (nsrc > 0) {
    DevBuf ds;
    ds.alloc(4 * (size_t)nsrc);
    GRE_HIP(hipMemcpyAsync(ds.p, srcs, 4 * (size_t)nsrc, hipMemcpyHostToDevice, s));
    dispatch(e->prog, [&](auto P) {
      hipLaunchKernelGGL(k_init_src<decltype(P)::value>, dim3(1), dim3(64), 0, s, c, ds.as<uint32_t>(), nsrc,
                         e->curr.as<uint32_t>());
    });
    GRE_HIP(hipStreamSynchronize(s));
  }
  GRE_HIP(hipGetLastError());
  const Adj a = make_adj(e);
  hipLaunchKernelGGL(k_count, dim3(e->nb), dim3(kB), 0, s, e->curr.as<uint32_t>(), e->g->nV, a,
                     e->bcnt.as<uint32_t>(), e->bdeg.as<uint64_t>());
  rebuild_queue(e, a, s);
  e->seen_edges = e->qedges;
  e->pulling = false;
  e->superstep = 0;
  e->pull_steps = 0;
  e->all_began = srcs == nullptr;
  e->begun = true;
  for (double& t : e->times) t = 0.0;
  e->times[0] = now() - t0;
}

// Source partition of the complete-frontier pull: 256 tail ranges and the
// 4,096 hottest sources shared (32 KB of FP64 scatter data, resident in every
// L2); used from 2^22 edges (below, the one-time partition build outweighs it)
// up to the partition's 32-bit entry limit.
constexpr int kEngRanges = 256;
constexpr uint32_t kEngShared = 1u << 12;
bool engine_part_ok(const gre_graph* g) { return g->nE >= (1ull << 22) && g->nE < 0xffffffffULL; }
// The binned FP64 pull (GRE's PageRank): the same edge range, and at most
// kBinMaxBins bins of kEngBinRows vertices with in-edges.
bool engine_bin_ok(gre_graph* g) {
  if (!engine_part_ok(g)) return false;
  graph_compact_targets(g);
  return (uint64_t)g->ntgt <= (uint64_t)kBinMaxBins * kEngBinRows;
}

int64_t engine_step(gre_engine* e) {
  GRE_CHECK(e->begun, GRE_DEV_NULL, "call gre_engine_begin first");
  hipStream_t s = stream();
  const Adj a = make_adj(e);
  EngCtx c = make_ctx(e);
  // phase times from events on the stream: no host wait between the phases
  for (hipEvent_t& ev : e->ev)
    if (!ev) GRE_HIP(hipEventCreate(&ev));
  GRE_HIP(hipEventRecord(e->ev[0], s));
  // nextSet.clear() (Engine.hpp:322)
  hipLaunchKernelGGL(k_zero_u32, dim3(gcap(e->nwords)), dim3(kB), 0, s, e->next.as<uint32_t>(), (uint64_t)e->nwords);
  // push while the frontier's edges are few, pull once they are an eighth of
  // the pull rows -- for sum combines (the push pays an atomic per edge, the
  // pull none) and first-writer ones (the pull stops at a row's first active
  // neighbour and skips reached vertices).  Min combines stay in the push
  // form: it skips the atomic wherever the label cannot fall, while the pull
  // reads every neighbour's label; on R-MAT scale 26 the pull step is slower
  // at every frontier size, the complete one included.
  const uint64_t epull = (e->dir == GRE_DIR_INOUT ? 2 : 1) * e->g->nE;
  const bool min_combine = e->prog == GRE_PROG_CC || e->prog == GRE_PROG_SSSP;
  // First-writer programs (BFS) reach each vertex once, so the edges of the
  // frontiers so far are the reached vertices' edges: they switch like a
  // direction-optimising BFS, to the pull once the frontier has a twentieth of
  // the edges not yet reached, back to the push once it holds under 1/24 of
  // the vertices.  Either form gives GRE's result; this only picks the cheaper.
  const bool first_writer = e->prog == GRE_PROG_BFS;
  bool auto_pull;
  if (first_writer) {
    const uint64_t unseen = epull > e->seen_edges ? epull - e->seen_edges : 0;
    auto_pull = e->pulling ? (uint64_t)e->qn * 24 > e->g->nV : e->qedges * 20 > unseen;
  } else {
    auto_pull = !min_combine && e->qedges * 8 > epull;
  }
  const bool pull = e->qedges > 0 && (e->mode == GRE_ENGINE_PULL || (e->mode == GRE_ENGINE_AUTO && auto_pull));
  e->pulling = pull;
  bool fused = false;   // the compute phase applied in place (kFuseApply)
  bool settled = false;  // ... and nothing can activate or halt: the active set and its counts stand
  const bool fsum = e->prog == GRE_PROG_PAGERANK || e->prog == GRE_PROG_PAGERANK_DELTA ||
                    e->prog == GRE_PROG_BC_BACKWARD;
  if (pull && e->prog == GRE_PROG_PAGERANK && e->dir == GRE_DIR_OUT && e->qedges == epull && engine_bin_ok(e->g)) {
    // GRE's PageRank with every vertex active: the binned pull.  The sources
    // are staged in compact order (their copy made once per run: GRE's
    // PageRank scatters its init value every superstep, PageRank.hpp:38-47),
    // each destination's sum is exact fixed point (binned.hip)
    gre_graph* g = e->g;
    if (!e->bxp) {
      graph_compact_sources(g);
      e->bxp = graph_binpart_engine(g);
    }
    if (e->bbuf.for_xp != e->bxp.get()) e->bbuf.make(*e->bxp, g->nsrc, s);
    // apply() fused into the pull: every vertex with in-edges gets a message
    // and runs apply(), which activates none (PageRank.hpp:70-74), so next
    // stays clear
    using PR = Prog<GRE_PROG_PAGERANK>;
    EngFuse f;
    f.X = (double*)c.X;
    f.a = PR::kResetProb;
    f.b = 1.0 - PR::kResetProb;
    bin_engine_sum(*e->bxp, e->bbuf, (const double*)c.S, g->cid.as<uint32_t>(), g->nV, !e->bsc_valid,
                   (double*)c.C, f, s);
    e->bsc_valid = true;
    fused = true;
    settled = true;   // PageRank never halts (kHalt = false) and its apply activates nothing
    ++e->pull_steps;
  } else if (pull && fsum && e->dir == GRE_DIR_OUT && e->qedges == epull && engine_part_ok(e->g)) {
    gre_graph* g = e->g;
    if (!e->xp) {
      graph_compact_sources(g);
      e->xp = graph_xpart(g, kEngRanges, kEngShared);
      e->seg_buf.make(e->seg_plan, *e->xp, 8, sizeof(double));
      e->seg_plan.nt = true;
      e->seg_plan.tag = "pull_compute";
      e->seg_plan.fix_tag = "pull_fixup";
      e->row_buf.make_pairs(e->row_plan, *e->xp, g->nV, 8, sizeof(double));
      e->row_plan.tag = "pull_combine";
      e->row_plan.fix_tag = "pull_combine_fixup";
      e->Sc.alloc(sizeof(double) * (size_t)g->nV);
      e->xsum.alloc(sizeof(double) * (size_t)e->xp->nseg);
    }
    const uint32_t nvpad = e->nwords * 32;
    // every vertex with a non-empty row gets a message: the same words every
    // superstep, computed once per engine
    if (!e->rows_mark.p) {
      e->rows_mark.alloc(4 * (size_t)e->nwords);
      hipLaunchKernelGGL(k_mark_rows, dim3(gcap(nvpad)), dim3(kB), 0, s, g->coff.as<uint64_t>(), g->nV, nvpad,
                         e->rows_mark.as<uint32_t>());
      GRE_HIP(hipGetLastError());
    }
    GRE_HIP(hipMemcpyAsync(e->next.p, e->rows_mark.p, 4 * (size_t)e->nwords, hipMemcpyDeviceToDevice, s));
    // GRE's PageRank scatters its init value every superstep (PageRank.hpp:38-47, apply leaves
    // pr_scatter alone), so its compact copy is made once per run
    if (!(e->sc_valid && e->prog == GRE_PROG_PAGERANK))
      hipLaunchKernelGGL(k_to_compact, dim3(gcap(g->nV)), dim3(kB), 0, s, (const double*)c.S, g->cid.as<uint32_t>(),
                         g->nV, e->Sc.as<double>());
    e->sc_valid = e->prog == GRE_PROG_PAGERANK;
    pull_pass(e->seg_plan, EngSegMode{e->Sc.as<double>(), e->xsum.as<double>()}, nullptr, s);
    pull_pass(e->row_plan, EngRowMode{e->xsum.as<double>(), (double*)c.C}, nullptr, s);
    ++e->pull_steps;
  } else if (pull && fsum && e->dir != GRE_DIR_INOUT && e->qedges == epull) {
    if (e->sum_dir != e->dir) {
      e->sum_buf.make(e->sum_plan, e->g, e->dir == GRE_DIR_OUT, 8, sizeof(double));
      e->sum_plan.tag = "pull_compute";
      e->sum_plan.fix_tag = "pull_fixup";
      e->sum_dir = e->dir;
    }
    const uint32_t nvpad = e->nwords * 32;
    hipLaunchKernelGGL(k_mark_rows, dim3(gcap(nvpad)), dim3(kB), 0, s, e->sum_plan.off, e->g->nV, nvpad,
                       e->next.as<uint32_t>());
    pull_pass(e->sum_plan, EngSumMode{(const double*)c.S, (double*)c.C}, nullptr, s);
    ++e->pull_steps;
  } else if (pull) {
    const PullAdj pa = make_pull_adj(e);
    const uint32_t nV = e->g->nV;
    dispatch(e->prog, [&](auto P) {
      constexpr int PV = decltype(P)::value;
      fused = Pull<PV>::kFuseApply;
      const PullRows& rows = ensure_pull_rows(e, pa, short_limit<Pull<PV>>(), s);
      ProfScope ps("pull_compute");
      // every neighbour's message is folded (no frontier test) when every
      // neighbour is active, or when the stale ones cannot change the result
      const bool all = e->qedges == epull || (Pull<PV>::kStale && e->all_began && e->qedges * 2 > epull);
      hipLaunchKernelGGL(k_pull_short<PV>, dim3(gcap(nV)), dim3(kB), 0, s, c, pa, nV, e->curr.as<uint32_t>(), all,
                         e->next.as<uint32_t>());
      if (rows.n_mid)
        hipLaunchKernelGGL((k_pull_rows<PV, 1>), dim3(std::min<uint32_t>((rows.n_mid + 3) / 4, 1u << 20)),
                           dim3(kB), 0, s, c, pa, rows.mid.as<uint32_t>(), rows.n_mid, e->curr.as<uint32_t>(), all,
                           e->next.as<uint32_t>());
      if (rows.n_long)
        hipLaunchKernelGGL((k_pull_rows<PV, 16>), dim3(std::min<uint32_t>(rows.n_long, 1u << 20)), dim3(1024), 0, s,
                           c, pa, rows.lng.as<uint32_t>(), rows.n_long, e->curr.as<uint32_t>(), all,
                           e->next.as<uint32_t>());
    });
    GRE_HIP(hipGetLastError());
    ++e->pull_steps;
  } else if (e->qedges > 0) {
    const uint64_t nblk = (e->qedges + kEB - 1) / kEB;
    GRE_CHECK(nblk < 0x7fffffffULL, GRE_DEV_FAILURE, "frontier too large for one launch");
    if (e->queue_stale) {
      hipLaunchKernelGGL(k_write, dim3(e->nb), dim3(kB), 0, s, e->curr.as<uint32_t>(), e->g->nV, a,
                         e->bcnt.as<uint32_t>(), e->bdeg.as<uint64_t>(), e->nb, e->queue.as<uint32_t>(),
                         e->qpos.as<uint64_t>());
      GRE_HIP(hipGetLastError());
      e->queue_stale = false;
    }
    ProfScope ps("push_expand");
    dispatch(e->prog, [&](auto P) {
      hipLaunchKernelGGL(k_expand<decltype(P)::value>, dim3((unsigned)nblk), dim3(kB), 0, s, c, a,
                         e->queue.as<uint32_t>(), e->qpos.as<uint64_t>(), e->qn, e->qedges, e->next.as<uint32_t>());
    });
    GRE_HIP(hipGetLastError());
  }
  GRE_HIP(hipEventRecord(e->ev[1], s));
  if (settled) {
    // currSet after the apply equals currSet before it, so the block counts,
    // the totals and the queue of the last rebuild still describe it
    GRE_HIP(hipEventRecord(e->ev[2], s));
    hipError_t q;
    while ((q = hipEventQuery(e->ev[2])) == hipErrorNotReady) {
    }
    GRE_HIP(q);
  } else {
    ProfScope ps("apply");
    dispatch(e->prog, [&](auto P) {
      if (fused)
        hipLaunchKernelGGL(k_activate<decltype(P)::value>, dim3(e->nb), dim3(kB), 0, s, e->next.as<uint32_t>(),
                           e->curr.as<uint32_t>(), a, e->bcnt.as<uint32_t>(), e->bdeg.as<uint64_t>());
      else
        hipLaunchKernelGGL(k_apply<decltype(P)::value>, dim3(e->nb), dim3(kB), 0, s, c, e->next.as<uint32_t>(),
                           e->curr.as<uint32_t>(), e->g->nV, a, e->bcnt.as<uint32_t>(), e->bdeg.as<uint64_t>());
    });
    GRE_HIP(hipGetLastError());
    rebuild_queue(e, a, s, e->ev[2]);   // waits for the stream
  }
  GRE_HIP(hipGetLastError());
  e->seen_edges += e->qedges;
  float ms_c = 0.f, ms_a = 0.f;
  GRE_HIP(hipEventElapsedTime(&ms_c, e->ev[0], e->ev[1]));
  GRE_HIP(hipEventElapsedTime(&ms_a, e->ev[1], e->ev[2]));
  e->times[2] += 1e-3 * ms_c;
  e->times[4] += 1e-3 * ms_a;
  // the program's reduce() hook: Forward increments currLevel (BC.hpp:40-42)
  if (e->prog == GRE_PROG_BC_FORWARD) e->level++;
  return e->qn;
}

}  // namespace
}  // namespace gre

using namespace gre;

extern "C" {

int gre_engine_create(gre_engine** out, gre_graph* g, int prog) {
  return guard([&] {
    GRE_CHECK(out && g, GRE_DEV_NULL, "null argument");
    auto* e = new gre_engine();
    try {
      e->g = g;
      e->prog = prog;
      size_t s = 0, c = 0, x = 0, aux = 0;
      switch (prog) {
        case GRE_PROG_PAGERANK: case GRE_PROG_PAGERANK_DELTA: s = c = x = 8; break;
        case GRE_PROG_BFS: case GRE_PROG_CC: s = c = x = 8; break;
        case GRE_PROG_SSSP: s = c = x = 4; e->c_is_x = true; break;
        case GRE_PROG_BC_FORWARD: s = 4; c = 4; x = 8; break;
        case GRE_PROG_BC_BACKWARD: s = c = x = 8; aux = 8; break;
        case GRE_PROG_COUNTIN: s = c = x = 4; break;
        default: throw Error(GRE_DEV_NULL, "unknown vertex program id");
      }
      const size_t n = g->nV;
      e->es[0] = s; e->es[1] = c; e->es[2] = x; e->es[3] = aux;
      e->S.alloc(s * n);
      if (!e->c_is_x) e->C.alloc(c * n);
      e->X.alloc(x * n);
      if (aux) e->AUX.alloc(aux * n);
      // bitmaps cover every vertex slot a kVB-vertex block touches, so the
      // apply/count/write blocks never index past them
      e->nb = (uint32_t)((n + kVB - 1) / kVB);
      e->nwords = e->nb * (uint32_t)(kVB / 32);
      e->curr.alloc(4 * (size_t)e->nwords);
      e->next.alloc(4 * (size_t)e->nwords);
      e->queue.alloc(4 * n);
      e->qpos.alloc(8 * (n + 1));
      e->bcnt.alloc(4 * ((size_t)e->nb + 1));
      e->bdeg.alloc(8 * ((size_t)e->nb + 1));
      default_state(e);
      GRE_HIP(hipStreamSynchronize(stream()));
    } catch (...) {
      delete e;
      throw;
    }
    *out = e;
  });
}

int gre_engine_destroy(gre_engine* e) {
  return guard([&] {
    if (!e) return;
    GRE_HIP(hipStreamSynchronize(stream()));
    // like a PageRank handle: drop the graph's cached partitions that nothing
    // holds once this engine is gone (about 9 GB each at 2^30 edges)
    gre_graph* g = e->g;
    for (hipEvent_t ev : e->ev)
      if (ev) (void)hipEventDestroy(ev);
    if (e->ev_tot) (void)hipEventDestroy(e->ev_tot);
    if (e->tot_host) (void)hipHostFree(e->tot_host);
    delete e;
    prune_xparts(g);
  });
}

int gre_engine_config_max_supersteps(gre_engine* e, int64_t n) {
  return guard([&] {
    GRE_CHECK(e, GRE_DEV_NULL, "null engine");
    e->max_ss = n > 0 ? n : 1024;
  });
}

int gre_engine_elem_size(const gre_engine* e, int which, size_t* bytes) {
  return guard([&] {
    GRE_CHECK(e && bytes, GRE_DEV_NULL, "null argument");
    GRE_CHECK(which >= 0 && which <= 3, GRE_DEV_NULL, "bad array id");
    *bytes = e->es[which];
  });
}

static gre::DevBuf& arr(gre_engine* e, int which) {
  GRE_CHECK(which >= 0 && which <= 3, GRE_DEV_NULL, "bad array id");
  GRE_CHECK(e->es[which] > 0, GRE_DEV_NULL, "program has no such array");
  if (which == GRE_ARR_COMBINE && e->c_is_x) return e->X;
  return which == 0 ? e->S : which == 1 ? e->C : which == 2 ? e->X : e->AUX;
}

int gre_engine_upload(gre_engine* e, int which, const void* host) {
  return guard([&] {
    GRE_CHECK(e && host, GRE_DEV_NULL, "null argument");
    gre::DevBuf& b = arr(e, which);
    GRE_HIP(hipMemcpyAsync(b.p, host, e->es[which] * (size_t)e->g->nV, hipMemcpyHostToDevice, stream()));
    GRE_HIP(hipStreamSynchronize(stream()));
    e->sc_valid = false;  // the scatter data may have changed
    e->bsc_valid = false;
  });
}

int gre_engine_download(gre_engine* e, int which, void* host) {
  return guard([&] {
    GRE_CHECK(e && host, GRE_DEV_NULL, "null argument");
    gre::DevBuf& b = arr(e, which);
    GRE_HIP(hipMemcpyAsync(host, b.p, e->es[which] * (size_t)e->g->nV, hipMemcpyDeviceToHost, stream()));
    GRE_HIP(hipStreamSynchronize(stream()));
  });
}

int gre_engine_set_level(gre_engine* e, int32_t level) {
  return guard([&] {
    GRE_CHECK(e, GRE_DEV_NULL, "null engine");
    e->level = level;
  });
}

int gre_engine_get_level(const gre_engine* e, int32_t* level) {
  return guard([&] {
    GRE_CHECK(e && level, GRE_DEV_NULL, "null argument");
    *level = e->level;
  });
}

int gre_engine_begin(gre_engine* e, int dir, const uint32_t* srcs, int64_t nsrc) {
  return guard([&] {
    GRE_CHECK(e, GRE_DEV_NULL, "null engine");
    engine_begin(e, dir, srcs, nsrc);
  });
}

int gre_engine_step(gre_engine* e, int64_t* active) {
  return guard([&] {
    GRE_CHECK(e, GRE_DEV_NULL, "null engine");
    const int64_t n = engine_step(e);
    if (active) *active = n;
  });
}

int gre_engine_run(gre_engine* e, int dir, const uint32_t* srcs, int64_t nsrc, int64_t* counts, int64_t counts_cap,
                   int64_t* supersteps) {
  return guard([&] {
    GRE_CHECK(e, GRE_DEV_NULL, "null engine");
    engine_begin(e, dir, srcs, nsrc);
    int64_t ss;
    for (ss = 0; ss < e->max_ss; ++ss) {
      const int64_t n = engine_step(e);
      if (counts && ss < counts_cap) counts[ss] = n;
      if (n == 0) break;
    }
    e->superstep = ss;
    if (supersteps) *supersteps = ss;
  });
}

int gre_engine_phase_times(const gre_engine* e, double* t5) {
  return guard([&] {
    GRE_CHECK(e && t5, GRE_DEV_NULL, "null argument");
    for (int i = 0; i < 5; ++i) t5[i] = e->times[i];
  });
}

int gre_engine_set_mode(gre_engine* e, int mode) {
  return guard([&] {
    GRE_CHECK(e, GRE_DEV_NULL, "null handle");
    GRE_CHECK(mode == GRE_ENGINE_AUTO || mode == GRE_ENGINE_PUSH || mode == GRE_ENGINE_PULL, GRE_DEV_NULL,
              "mode must be GRE_ENGINE_AUTO, _PUSH or _PULL");
    e->mode = mode;
  });
}

int gre_engine_pull_steps(const gre_engine* e, int64_t* steps) {
  return guard([&] {
    GRE_CHECK(e && steps, GRE_DEV_NULL, "null argument");
    *steps = e->pull_steps;
  });
}

}  // extern "C"
