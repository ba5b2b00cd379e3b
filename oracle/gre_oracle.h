/*
 * gre_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of GRE's scatter -> combine -> apply superstep and the
 * textbook algorithms that pin it.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library, and only as the checker (or
 * as the timed CPU baseline).  The product path (graphine_amd/libgre_dev.so)
 * never links, loads or calls it.
 *
 * Reference files restated here (paths relative to the GRE source tree):
 *   include/Engine/Engine.hpp:246-406,422-814,1377-1389  superstep driver
 *   include/Graph/VertexSet.hpp:17-284, include/Util/Bitmap.hpp:12-130
 *   include/Graph/LocalGraph.hpp:66-195                   CSR / CSC build
 *   include/Graph/GraphIngress.hpp:76-304,511-723         np=1 edge order
 *   include/Kernels/{PageRank,BFS,CC,SSSP,BC,CountIn}.hpp vertex programs
 */
#ifndef GRE_ORACLE_H
#define GRE_ORACLE_H
#include <stdint.h>
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- deterministic R-MAT (same definition as the device generator) ---- */
uint64_t o_mix64(uint64_t k);
/* Generates ef<<scale edges, drops self-loops, optionally symmetrises,
 * sorts by (src,dst) and removes duplicates.  keys must hold
 * (ef<<scale)*(sym?2:1) entries; returns the unique count; key = src<<32|dst. */
size_t o_rmat_keys(int scale, int ef, uint64_t seed, int flags, uint64_t* keys);
uint32_t o_rmat_scramble(int scale, uint64_t seed, uint32_t x);
void o_rmat_scramble_many(int scale, uint64_t seed, uint64_t n, const uint32_t* x, uint32_t* y);
/* Raw draws first..first+n-1 (self-loops and duplicates kept). */
void o_rmat_raw(int scale, uint64_t seed, uint64_t first, uint64_t n, uint32_t* u, uint32_t* v);
void o_sort_unique_keys(uint64_t* keys, size_t* n);
/* CSR from sorted keys: off[nV+1], idx[nE]. */
void o_csr_from_keys(uint32_t nV, size_t nE, const uint64_t* keys, uint64_t* off, uint32_t* idx);
/* CSR from an edge list in input order (GRE np=1 ingress: stable in-row order). */
void o_csr_from_edges(uint32_t nV, size_t nE, const uint32_t* src, const uint32_t* dst,
                      uint64_t* off, uint32_t* idx);
/* CSC from CSR, stable (in-row order = CSR position order); ceid = CSR position. */
void o_csc_from_csr(uint32_t nV, const uint64_t* off, const uint32_t* idx,
                    uint64_t* coff, uint32_t* cidx, uint64_t* ceid);
/* Edge weights. */
uint32_t o_sssp_weight(uint64_t seed, uint32_t u, uint32_t v);
void o_sssp_weights_csr(uint64_t seed, uint32_t nV, const uint64_t* off, const uint32_t* idx, uint32_t* w);
double o_lognormal_weight(uint64_t seed, uint32_t u, uint32_t v);

/* ---- GRE engine restatement ---- */
enum { O_PAGERANK = 0, O_PAGERANK_DELTA = 1, O_BFS = 2, O_CC = 3, O_SSSP = 4,
       O_BC_FORWARD = 5, O_BC_BACKWARD = 6, O_COUNTIN = 7 };
enum { O_DIR_OUT = 0, O_DIR_IN = 1, O_DIR_INOUT = 2 };

typedef struct {
  uint32_t nV;
  const uint64_t* off; const uint32_t* idx;                     /* CSR */
  const uint64_t* coff; const uint32_t* cidx; const uint64_t* ceid; /* CSC */
  const int32_t* outdeg;
} o_graph;

/* Runs SynchronousEngine<prog>::run<dir>() on one process.
 * S/C/X are the program's scatter/combine/vertex arrays, initialised by the
 * caller exactly as the program's alloc() would.  aux: SSSP edge weights in
 * CSR order (uint32), or BC_BACKWARD's forward state ({int level,int numPaths}).
 * level_io: BC currLevel (read and written back).  srcs==NULL means setALL().
 * counts[k] receives the active count after superstep k.
 * Returns getSuperstepCount(). */
int64_t o_gre_run(int prog, int dir, const o_graph* g, void* S, void* C, void* X,
                  const void* aux, int32_t* level_io, const uint32_t* srcs, int64_t nsrc,
                  int64_t max_ss, int64_t* counts, int64_t counts_cap);

/* ---- textbook references ---- */
void o_bfs_levels(uint32_t nV, const uint64_t* off, const uint32_t* idx, uint32_t src, int32_t* level);
void o_cc_unionfind(uint32_t nV, const uint64_t* off, const uint32_t* idx, uint32_t* label);
void o_dijkstra(uint32_t nV, const uint64_t* off, const uint32_t* idx, const uint32_t* w,
                uint32_t src, uint32_t* dist);
/* Brandes for one source: sigma (path counts), delta (dependencies), level. */
void o_brandes_one(uint32_t nV, const uint64_t* off, const uint32_t* idx, uint32_t src,
                   double* sigma, double* delta, int32_t* level);
/* FP64 PageRank: 1/N init, uniform teleport, dangling mass redistributed uniformly. */
void o_pagerank(uint32_t nV, const uint64_t* coff, const uint32_t* cidx, const int32_t* outdeg,
                double damping, int iters, double* x);
/* Weighted personalised PageRank (FP64).  wc: normalised weights in CSC order.
 * dangling mass is redistributed along p.  Iterates until L1 change < tol.
 * Returns iterations done. */
int o_ppr(uint32_t nV, const uint64_t* coff, const uint32_t* cidx, const double* wc,
          const int32_t* outdeg, const double* p, double damping, double tol, int maxit, double* x);
/* Adjoint: lambda = c + d*(W lambda + g (p.lambda)), W in CSR order (wr). */
int o_ppr_adjoint(uint32_t nV, const uint64_t* off, const uint32_t* idx, const double* wr,
                  const int32_t* outdeg, const double* p, const double* c, double damping,
                  double tol, int maxit, double* lam);
/* dL/dW_uv = d*x_u*lambda_v ; dL/dw_raw via the row normalisation. CSR order. */
void o_ppr_edge_grad(uint32_t nV, const uint64_t* off, const uint32_t* idx, const double* wr,
                     const double* rowsum, const double* x, const double* lam, double damping,
                     double* gW, double* graw);

/* ---- GRE-style multi-threaded push superstep (CPU baseline timing) ---- */
/* PageRank power iteration in GRE's scatter/combine/apply form: 12 workers
 * (include/Engine/Engine.hpp:211) push x_u/deg_u along CSR edges into a
 * combine array under 8192 striped spin-locks (include/Util/vLock.hpp:75-99),
 * then apply.  Returns seconds spent in the supersteps. */
double o_gre_pagerank_mt(int nthreads, uint32_t nV, const uint64_t* off, const uint32_t* idx,
                         const int32_t* outdeg, double damping, int iters, double* x);
/* vertex-cut ingress at np processes */
int o_ldg_assign(uint32_t nV, uint64_t nE_meta, int np, int nfiles, const uint64_t* file_off,
                 const uint32_t* src, const uint32_t* dst, int32_t* part);
void o_vertex_cut_stats(uint32_t nV, int np, int dir, uint64_t nE, const uint32_t* src, const uint32_t* dst,
                        const int32_t* part, int64_t* stats);
/* PageRank supersteps at np processes of W workers each on one host, over the
 * LDG vertex cut (CPU baseline timing; see gre_oracle.c) */
int o_gre_pagerank_np(int np, int nw, const int* workers, uint32_t nV, const uint64_t* off, const uint32_t* idx,
                      const int32_t* outdeg, double damping, int iters, double* x, double* secs, double* info);

#ifdef __cplusplus
}
#endif

#endif
