/*
 * gre_dev.h -- C ABI of the MI355X (gfx950) superstep library libgre_dev.so.
 *
 * This is the drop-in boundary for GRE's scatter -> combine -> apply superstep
 * (include/Engine/Engine.hpp:246-814 in the GRE tree).  GRE itself exposes a
 * header-only C++ template API and no C ABI; the entry points below are what
 * that API's hot path needs from a device, and the C++ host templates in
 * graphine_amd/include (SynchronousEngine<VC>::run<dir>() and friends) are the
 * only callers a GRE user program sees.  Each declaration names the GRE
 * interface it replaces.
 *
 * Conventions
 *   - every function returns an int status with GRE_Status values
 *     (include/Status.hpp:5-15): 0 = GRE_SUCCESS, 1 = GRE_FAILURE,
 *     3 = GRE_NULL (bad argument), 8 = GRE_NotImplemented;
 *     gre_last_error() returns a message for the calling thread;
 *   - host pointers unless a name says _dev; no C++ types or exceptions cross;
 *   - all work for one handle is ordered on the library stream
 *     (gre_set_stream); one host thread per handle;
 *   - vertex ids are uint32, edge offsets uint64 (entries can exceed 2^31).
 */
#ifndef GRE_DEV_H
#define GRE_DEV_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GRE_DEV_SUCCESS = 0, GRE_DEV_FAILURE = 1, GRE_DEV_NOBUFFER = 2, GRE_DEV_NULL = 3,
  GRE_DEV_NOT_IMPLEMENTED = 8
};

/* GRE's vertex programs (the headers under include/Kernels). */
enum {
  GRE_PROG_PAGERANK = 0,        /* include/Kernels/PageRank.hpp:18-126  */
  GRE_PROG_PAGERANK_DELTA = 1,  /* include/Kernels/PageRank.hpp:129-234 */
  GRE_PROG_BFS = 2,             /* include/Kernels/BFS.hpp:18-86        */
  GRE_PROG_CC = 3,              /* include/Kernels/CC.hpp:18-125        */
  GRE_PROG_SSSP = 4,            /* include/Kernels/SSSP.hpp:19-96       */
  GRE_PROG_BC_FORWARD = 5,      /* include/Kernels/BC.hpp:22-95         */
  GRE_PROG_BC_BACKWARD = 6,     /* include/Kernels/BC.hpp:97-168        */
  GRE_PROG_COUNTIN = 7          /* include/Kernels/CountIn.hpp:18-106   */
};
/* enum Dir{Out, In, InOut} (include/Types.hpp:31) */
enum { GRE_DIR_OUT = 0, GRE_DIR_IN = 1, GRE_DIR_INOUT = 2 };
/* state arrays of a vertex program (VertexCode.hpp:117-119) */
enum { GRE_ARR_SCATTER = 0, GRE_ARR_COMBINE = 1, GRE_ARR_VERTEX = 2, GRE_ARR_AUX = 3 };
enum { GRE_F32 = 0, GRE_F64 = 1 };

/* ---- library ----------------------------------------------------------- */
const char* gre_last_error(void);
/* Selects the device and checks it is gfx950.  No GRE counterpart
 * (COMM::Underlying::init, include/Comm/CommUnderlying.hpp:36-52, is the
 * closest: it brings the process up). */
int gre_init(int device);
/* All subsequent launches go to this hipStream_t (NULL = library stream). */
int gre_set_stream(void* stream);
int gre_get_stream(void** stream);
int gre_sync(void);
/* Per-kernel HIP-event timing of the library's hot kernels (off by default). */
int gre_prof_enable(int on);
int gre_prof_reset(void);
/* name: "pull_tile", "push_expand", ...; returns launches and total ms */
int gre_prof_get(const char* name, int64_t* launches, double* total_ms);
/* Device memory the library's handles and scratch hold, in bytes. */
int gre_mem_in_use(uint64_t* bytes);
/* Measurement aid (no GRE counterpart): times a plain streaming copy of
 * `bytes` bytes, reps times, on the library stream; the kernel that calibrates
 * HBM byte counters and gives the achievable copy bandwidth. */
int gre_copy_probe(uint64_t bytes, int reps, double* ms_per_copy);
/* Measurement aid (no GRE counterpart): times a read-only stream over `bytes`
 * bytes (16-byte loads, eight in flight per thread), reps times: the
 * achievable HBM read bandwidth. */
int gre_read_probe(uint64_t bytes, int reps, double* ms_per_read);
/* Measurement aid (no GRE counterpart): times `gathers` uniform random 4-byte
 * gathers from a table of table_entries floats (ids streamed from HBM), reps
 * times: the ceiling of scattered reads served from L1, L2, the Infinity
 * Cache or HBM, depending on the table size. */
int gre_gather_probe(uint64_t table_entries, uint64_t gathers, int reps, double* ms_per_pass);

/* ---- graph: device-resident CSR + CSC ------------------------------------
 * Replaces DistributedGraph<GraphInCSR>::ingress + GraphInCSR::build/rbuild
 * (include/Graph/DistributedGraph.hpp:136-172, include/Graph/LocalGraph.hpp:66-156)
 * at one process: lvid == vid, masters = all vertices. */
typedef struct gre_graph gre_graph;
/* From an edge list in input order: CSR rows keep input order (GraphInCSR::build),
 * CSC rows keep input order (GraphInCSR::rbuild). */
int gre_graph_from_edges(gre_graph** g, uint32_t nV, uint64_t nE, const uint32_t* src,
                         const uint32_t* dst);
/* From a host CSR (off[nV+1], idx[nE]); CSC is built on the device, stable. */
int gre_graph_from_csr(gre_graph** g, uint32_t nV, uint64_t nE, const uint64_t* off,
                       const uint32_t* idx);
/* Seeded R-MAT generated and built on the device: ef<<scale draws,
 * (a,b,c,d)=(0.57,0.19,0.19,0.05), self-loops and duplicates removed,
 * rows sorted.  flags: GRE_RMAT_SYM adds the reverse of every edge;
 * GRE_RMAT_SCRAMBLE permutes the vertex labels by a seeded bijection (as
 * Graph500's generator does), so the hubs are not the low ids. */
enum { GRE_RMAT_SYM = 1, GRE_RMAT_SCRAMBLE = 2 };
int gre_graph_rmat(gre_graph** g, int scale, int ef, uint64_t seed, int flags);
int gre_graph_destroy(gre_graph* g);
int gre_graph_info(const gre_graph* g, uint32_t* nV, uint64_t* nE);
/* Any output may be NULL.  off/coff: nV+1 uint64; idx/cidx: nE uint32;
 * ceid: nE uint64 (CSR position of each CSC entry); outdeg: nV int32. */
int gre_graph_download(const gre_graph* g, uint64_t* off, uint32_t* idx, uint64_t* coff,
                       uint32_t* cidx, uint64_t* ceid, int32_t* outdeg);
/* Per-edge uint32 weights for SSSP, in CSR order (the file order GRE's
 * DistributedGraph::loadEdgeProperty reads, DistributedGraph.hpp:194-206).
 * w == NULL: generate 1 + hash(seed, min(u,v), max(u,v)) % 255 on the device. */
int gre_graph_set_weights_u32(gre_graph* g, const uint32_t* w, uint64_t seed);
int gre_graph_get_weights_u32(const gre_graph* g, uint32_t* w);

/* ---- GRE engine: SynchronousEngine<VC>::run<dir>() on the device ---------
 * (include/Engine/Engine.hpp:53-1418).  One handle = one engine + the vertex
 * program's scatter/combine/vertex arrays in HBM. */
typedef struct gre_engine gre_engine;
/* Arrays are allocated and filled as the program's alloc() leaves them. */
int gre_engine_create(gre_engine** e, gre_graph* g, int prog);
int gre_engine_destroy(gre_engine* e);
/* configMaxSupersteps(n) (Engine.hpp:148-153): n<=0 restores 1024. */
int gre_engine_config_max_supersteps(gre_engine* e, int64_t n);
/* Bytes of one element of array `which` (GRE_ARR_*); 0 if absent. */
int gre_engine_elem_size(const gre_engine* e, int which, size_t* bytes);
/* Host <-> device copies of a program array (nV elements; AUX = BC_BACKWARD's
 * forward state, nV x {int32 level, int32 numPaths}). */
int gre_engine_upload(gre_engine* e, int which, const void* host);
int gre_engine_download(gre_engine* e, int which, void* host);
/* Forward's/Backward's currLevel (BC.hpp:45-46,106-107). */
int gre_engine_set_level(gre_engine* e, int32_t level);
int gre_engine_get_level(const gre_engine* e, int32_t* level);
/* run<dir>(srcSet): srcs == NULL means srcSet.setALL() (VertexSet.hpp:307).
 * counts (optional) receives the active-vertex count of each superstep, the
 * value GRE prints as "Superstep-k: Number of active vertices" (Engine.hpp:396-398).
 * supersteps receives getSuperstepCount() (Engine.hpp:1307-1309). */
int gre_engine_run(gre_engine* e, int dir, const uint32_t* srcs, int64_t nsrc,
                   int64_t* counts, int64_t counts_cap, int64_t* supersteps);
/* Stepwise form of run(): begin = init phase, step = one superstep
 * (compute + apply + vote), returning the active count. */
int gre_engine_begin(gre_engine* e, int dir, const uint32_t* srcs, int64_t nsrc);
int gre_engine_step(gre_engine* e, int64_t* active);
/* Phase times of the last run in seconds: init, sync-scatter, compute,
 * sync-combiner, apply (Engine.hpp:1292-1306; the two sync phases are 0). */
int gre_engine_phase_times(const gre_engine* e, double* t5);
/* Compute-phase form: AUTO (default) pushes along the frontier's edges while
 * they are under 1/8 of the direction's edges and otherwise pulls (each
 * destination folds its active neighbours' messages with the program's
 * combine, no atomics); CC and SSSP (min combines) always push under AUTO.
 * PUSH and PULL force one form.  GRE has only the push form
 * (Engine.hpp:592-625); both give GRE's results. */
enum { GRE_ENGINE_AUTO = 0, GRE_ENGINE_PUSH = 1, GRE_ENGINE_PULL = 2 };
int gre_engine_set_mode(gre_engine* e, int mode);
/* Supersteps of the last run that used the pull form. */
int gre_engine_pull_steps(const gre_engine* e, int64_t* steps);

/* ---- PageRank / personalised PageRank (pull over CSC) --------------------
 * No GRE counterpart: GRE's shipped PageRank never refreshes its scatter data
 * (PageRank.hpp:70-74).  This is the standard power iteration the PageRank
 * vertex-program model expresses: 1/N start, uniform teleport, dangling mass
 * redistributed uniformly (or along p for personalised runs). */
typedef struct gre_pr gre_pr;
/* k = number of vectors (batched PPR: row-major nV x k), dtype GRE_F32/GRE_F64,
 * weighted = 1: per-edge weights (set with gre_pr_set_edge_weights). */
int gre_pr_create(gre_pr** p, gre_graph* g, int dtype, int k, int weighted);
int gre_pr_destroy(gre_pr* p);
/* Raw per-edge weights in CSR order (FP32, nE); the library row-normalises.
 * w == NULL: lognormal(0,1) from a hash of (seed, u, v). */
int gre_pr_set_edge_weights(gre_pr* p, const float* w, uint64_t seed);
int gre_pr_get_edge_weights(gre_pr* p, float* w_raw, float* w_norm);
/* Personalisation (nV x k, row-major, each column sums to 1); NULL = uniform. */
int gre_pr_set_personalization(gre_pr* p, const void* pvec);
/* Runs until iters iterations are done or the FP64 L1 change of every vector
 * is below tol (tol <= 0: fixed iters).  Returns iterations and residual. */
int gre_pr_run(gre_pr* p, double damping, int iters, double tol, int* iters_done,
               double* residual);
/* Single-vector runs with tol <= 0 (a fixed iteration count) are captured
 * once into a hipGraph and replayed while the handle's schedule, buffers and
 * parameters are unchanged (on = 1, the default); on = 0 launches eagerly. */
int gre_pr_set_graphs(gre_pr* p, int on);
/* No counterpart in GRE: the same single-vector forward as gre_pr_run, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It calls no
 * hipMalloc/hipFree or allocating library routine: its reductions live in a
 * static 1.06 MB device array per GPU, and it launches eagerly (no captured
 * graph of its own).  Limits: k = 1 (GRE_DEV_NOT_IMPLEMENTED otherwise),
 * iters <= 4096 (GRE_DEV_NOBUFFER), kernel profiling off, and the handle's pull
 * plan already built for its current schedule, partition and weights by an
 * earlier gre_pr_run (GRE_DEV_NOBUFFER otherwise).  With tol <= 0 and
 * residual == NULL it only enqueues the work on the library stream (no host
 * wait).  Calls that overlap on one GPU share the array and race. */
int gre_pr_run_noalloc(gre_pr* p, double damping, int iters, double tol, int* iters_done,
                       double* residual);
int gre_pr_get_ranks(gre_pr* p, void* host_out);
/* Adjoint of the converged forward: lambda = c + d (W lambda + g (p.lambda)),
 * a pull over CSR; c = indicator of targets (ntgt ids, or dense nV x k in cvec). */
int gre_pr_adjoint(gre_pr* p, const uint32_t* targets, int64_t ntgt, double damping, int iters,
                   double tol, int* iters_done, double* residual);
/* The same adjoint seeded with an arbitrary cotangent c = d(loss)/d(ranks)
 * (host, nV x k in the handle's dtype, the layout of gre_pr_get_ranks). */
int gre_pr_adjoint_cotangent(gre_pr* p, const void* cot, double damping, int iters, double tol,
                             int* iters_done, double* residual);
/* d(loss)/d(weights): gW = d x_u lambda_v (normalised weights), graw via the
 * row normalisation; CSR order, summed over the k vectors; d(loss)/d(p). */
int gre_pr_gradients(gre_pr* p, double damping, float* gW, float* graw, void* gp);
int gre_pr_get_adjoint(gre_pr* p, void* host_out);
/* Pull schedules of the forward pass.  All are tile-segmented reductions with
 * tile_items (4, 8 or 16) adjacency entries per thread; _NT streams the
 * neighbour ids with a non-temporal hint; COMPACT gathers from a contrib array
 * indexed by hub-sorted compact source ids (single vector; weighted handles
 * gather the ranks there and multiply by the edge weight).
 * PART additionally cuts the compact sources into ranges, each pinned to one
 * XCD so its gathers stay in that XCD's L2, and sums a row's per-range partial
 * sums in a second pull (single vector, fewer than 2^32 edges; ranges set with
 * gre_pr_set_partition; weighted handles keep a copy of the weights in the
 * partition's entry order, remade when the weights or the partition change).
 * Under COMPACT and PART the single-vector adjoint pulls over the CSR in
 * compact target ids (vertices with in-edges by descending in-degree); under
 * PART, partitioned over the CSR with the handle's ranges.  Under BIN it runs
 * the binned pull over the reversed graph (fixed-point scale set per iteration
 * from max |lambda|); a graph with more than 4096 x 16384 vertices with
 * out-edges falls back to PART's adjoint with the ranges of the best
 * partition the autotuner timed, or the defaults.
 * Batched handles (k > 1) have two schedules: TILE (groups of lanes walk
 * fixed tiles of entries, rows split across tiles are joined by a fix-up) and
 * ROWS (one group of lanes per destination row, wavefront per vertex).
 * BIN (single vector, FP32, fewer than 2^32 edges, personalisation of L1 norm
 * at most 1) is a propagation-blocking pull: chunks of source rows stage their
 * values in LDS, sum each destination row's entries of the chunk (at most 16
 * at a time, in a fixed order) and stream the sums into per-bin runs
 * (weighted handles multiply by the normalised weights, kept in the layout's
 * entry order), then each bin of destination rows adds its sums in LDS as
 * 64-bit fixed point (exact, so the result does not depend on the order of
 * the additions); its rows are relabelled like PART's.
 * GRE_SCHED_AUTO times every candidate on the first run and keeps the fastest
 * (for batched handles, per direction: forward and adjoint). */
enum { GRE_SCHED_AUTO = 0, GRE_SCHED_TILE = 1, GRE_SCHED_TILE_NT = 2, GRE_SCHED_COMPACT = 3,
       GRE_SCHED_COMPACT_NT = 4, GRE_SCHED_PART = 5, GRE_SCHED_PART_NT = 6, GRE_SCHED_ROWS = 7,
       GRE_SCHED_BIN = 8 };
int gre_pr_set_schedule(gre_pr* p, int schedule, int tile_items);
int gre_pr_get_schedule(gre_pr* p, int* schedule, int* tile_items);
/* PART schedules: the shared_sources hottest compact sources (0 = none; fewer
 * if the graph has fewer sources) form one range pulled by all XCDs, the rest
 * are cut into `ranges` tail ranges (8, 16, 32, 64, 128 or 256) pinned to
 * XCDs.  Defaults 64 and 2^18; tuned. */
int gre_pr_set_partition(gre_pr* p, int ranges, uint32_t shared_sources);
int gre_pr_get_partition(gre_pr* p, int* ranges, uint32_t* shared_sources);

/* ---- traversal kernels (plain single-source paths) ------------------------ */
/* Direction-optimising BFS levels (int32, -1 unreached). edges_out: sum of the
 * degrees of reached vertices. */
int gre_bfs(gre_graph* g, uint32_t src, int32_t* level_out, uint64_t* edges_out,
            int32_t* depth_out);
/* Same, output left on the device (for timing); level_dev may be NULL. */
int gre_bfs_dev(gre_graph* g, uint32_t src, int32_t* level_dev, uint64_t* edges_out,
                int32_t* depth_out);
/* Multi-source BFS: up to 64 sources as a 64-bit mask per vertex.
 * level_out (optional): nsrc x nV int32 (source-major). */
int gre_msbfs(gre_graph* g, const uint32_t* srcs, int nsrc, int32_t* level_out,
              uint64_t* edges_out);
/* Brandes BC for nsrc sources in one launch sequence, FP64 sigma/delta.
 * bc_out: nV doubles (sum of delta over sources, source excluded). Optional
 * sigma_out/delta_out: nsrc x nV. */
int gre_bc(gre_graph* g, const uint32_t* srcs, int nsrc, double* bc_out, double* sigma_out,
           double* delta_out, int32_t* level_out);
/* Long-row chunking of the batched traversals, per graph (no GRE counterpart:
 * GRE's engine has no such kernels; the knob is to these what
 * gre_pr_set_schedule is to PageRank).  which: GRE_CHUNKS_MSBFS (the in-rows
 * of gre_msbfs's bottom-up levels) or GRE_CHUNKS_BC_FORWARD (gre_bc's forward
 * pull).  Rows with more than long_rows entries are cut into chunks of
 * `chunk` entries, one wavefront each.  long_rows = 0 (the default) leaves the
 * pair to the first call on the graph with at least 32 sources, which times a
 * few candidates on them and keeps the fastest (narrower calls before it run
 * a fixed default pair).  Levels and path counts do not depend on the pair; the BC
 * backward pass keeps a fixed one (its partial sums set the rounding of the
 * dependencies). */
enum { GRE_CHUNKS_MSBFS = 0, GRE_CHUNKS_BC_FORWARD = 1 };
int gre_traversal_set_chunks(gre_graph* g, int which, uint64_t long_rows, uint64_t chunk);
int gre_traversal_get_chunks(gre_graph* g, int which, uint64_t* long_rows, uint64_t* chunk);

/* ---- vertex-cut placement for np processes (GRE's MPI ingress) ------------
 * Host-side, no device work: the placement a run over np GPUs would load.
 * Replaces GraphIngress::LDGEdgeDecision (include/Graph/GraphIngress.hpp:
 * 1126-1171) as loadEdgesFromFiles drives it (:511-723): process r reads the
 * edge files r, r + np, ... in order and decides each edge's part with its own
 * state; capacity = nE_meta / np / np (must be > 0, i.e. nE_meta >= np*np,
 * where GRE would divide by zero).  file_off[nfiles + 1] cuts src/dst (all
 * files concatenated, in file order) into the files; part receives each
 * edge's part. */
int gre_vertex_cut_ldg(uint32_t nV, uint64_t nE_meta, int np, int nfiles, const uint64_t* file_off,
                       const uint32_t* src, const uint32_t* dst, int32_t* part);
/* Per part p (stats[7 p ...]): masters, scatters, combiners, bi-agents,
 * agents, local edges, aligned masters, as GraphIngress::ingress<dir> counts
 * them (:111-153, 229-235); owner(v) = v % np.  The PageRank program prints
 * the sums of combiners and scatters and their ratio to the edges
 * (include/Kernels/PageRank.hpp:103-122). */
int gre_vertex_cut_stats(uint32_t nV, int np, int dir, uint64_t nE, const uint32_t* src, const uint32_t* dst,
                         const int32_t* part, int64_t* stats);
/* Part p's lvid2vid: masters p, p+np, ..., padding to a multiple of 64 with
 * all-ones, then its combiners (Out) or all agents (In, InOut) in ascending
 * id (:154-178).  n_out receives the length; lvid2vid may be NULL. */
int gre_vertex_cut_lvid2vid(uint32_t nV, int np, int dir, int p, uint64_t nE, const uint32_t* src,
                            const uint32_t* dst, const int32_t* part, uint64_t* lvid2vid, uint64_t cap,
                            uint64_t* n_out);

#ifdef __cplusplus
}
#endif
#endif
