/*
 * gre_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A CPU restatement of GRE's superstep and of the textbook algorithms that pin
 * it.  tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg are the
 * only users; the product library never links or calls anything here.
 *
 * What pins each part (checked in tests/test_oracle_*.py):
 *   R-MAT generator        fixture tests/golden/rmat_*.npz + quadrant statistics
 *   CSR/CSC build          brute-force edge-set comparison, stable order checks
 *   GRE engine, BFS        o_bfs_levels (textbook) level sizes == active counts
 *   GRE engine, CC         o_cc_unionfind (different algorithm), min-vid labels
 *   GRE engine, SSSP       o_dijkstra (different algorithm)
 *   GRE engine, BC         o_brandes_one and brute-force all-pairs path counts
 *   GRE engine, PageRank   closed form pr = 0.15 + 0.85*sum(0.15/outdeg)
 *   GRE engine, CountIn    in-degree counts, sum(in) == sum(out)
 *   o_pagerank / o_ppr     dense Google-matrix iteration and linear solve (numpy)
 *   o_ppr_adjoint / grad   numpy linear solve of the transposed system and
 *                          central finite differences of an FP64 forward
 *   GRE engine, PageRank_delta  numpy matrix powers, the linear-solve fixpoint
 *                          of a run to halt, a hand-built 1e-5 knife edge
 *   o_gre_pagerank_mt/_np  textbook o_pagerank ranks; the np run's message
 *                          counts equal o_vertex_cut_stats' agent counts
 */
#define _POSIX_C_SOURCE 200809L
#include "gre_oracle.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <time.h>
#include <pthread.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Hashing and R-MAT                                                        */
/* ------------------------------------------------------------------------ */
#define K0 0x9E3779B97F4A7C15ULL
#define K1 0xBF58476D1CE4E5B9ULL
#define K2 0x94D049BB133111EBULL
#define K3 0x2545F4914F6CDD1DULL
#define K4 0xD6E8FEB86659FD93ULL

uint64_t o_mix64(uint64_t k) {
  /* murmur3 fmix64 finaliser */
  k ^= k >> 33; k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

/* quadrant thresholds for (a,b,c,d) = (0.57,0.19,0.19,0.05) on 32-bit draws */
#define RMAT_TA   ((57ULL << 32) / 100)
#define RMAT_TAB  ((76ULL << 32) / 100)
#define RMAT_TABC ((95ULL << 32) / 100)

static inline void rmat_edge(int scale, uint64_t seed, uint64_t i, uint32_t* pu, uint32_t* pv) {
  uint32_t u = 0, v = 0;
  uint64_t h = 0;
  for (int l = 0; l < scale; ++l) {
    if ((l & 1) == 0) h = o_mix64(seed * K0 + i * K1 + (uint64_t)(l >> 1) * K2);
    const uint64_t r = (l & 1) ? (h & 0xffffffffULL) : (h >> 32);
    uint32_t ub, vb;
    if (r < RMAT_TA) { ub = 0; vb = 0; }
    else if (r < RMAT_TAB) { ub = 0; vb = 1; }
    else if (r < RMAT_TABC) { ub = 1; vb = 0; }
    else { ub = 1; vb = 1; }
    u = (u << 1) | ub;
    v = (v << 1) | vb;
  }
  *pu = u; *pv = v;
}

static void radix_sort_u64(uint64_t* a, uint64_t* tmp, size_t n) {
  uint64_t* src = a; uint64_t* dst = tmp;
  size_t cnt[256];
  for (int pass = 0; pass < 8; ++pass) {
    const int sh = pass * 8;
    memset(cnt, 0, sizeof(cnt));
    for (size_t i = 0; i < n; ++i) cnt[(src[i] >> sh) & 255]++;
    int trivial = 0;
    for (int b = 0; b < 256; ++b) if (cnt[b] == n) trivial = 1;
    if (trivial) continue;
    size_t s = 0;
    for (int b = 0; b < 256; ++b) { size_t c = cnt[b]; cnt[b] = s; s += c; }
    for (size_t i = 0; i < n; ++i) dst[cnt[(src[i] >> sh) & 255]++] = src[i];
    uint64_t* t = src; src = dst; dst = t;
  }
  if (src != a) memcpy(a, src, n * sizeof(uint64_t));
}

void o_sort_unique_keys(uint64_t* keys, size_t* n) {
  size_t m = *n;
  if (m == 0) return;
  uint64_t* tmp = (uint64_t*)malloc(m * sizeof(uint64_t));
  radix_sort_u64(keys, tmp, m);
  free(tmp);
  size_t w = 1;
  for (size_t i = 1; i < m; ++i)
    if (keys[i] != keys[w - 1]) keys[w++] = keys[i];
  *n = w;
}

void o_rmat_raw(int scale, uint64_t seed, uint64_t first, uint64_t n, uint32_t* u, uint32_t* v) {
  for (uint64_t i = 0; i < n; ++i) rmat_edge(scale, seed, first + i, &u[i], &v[i]);
}

/* Label scramble (as Graph500's generator permutes its vertex labels): a
 * bijection of [0, 2^scale) keyed by the seed, built from steps that are each
 * invertible modulo 2^scale (add, multiply by an odd number, xor with a right
 * shift of itself). */
uint32_t o_rmat_scramble(int scale, uint64_t seed, uint32_t x) {
  const uint64_t mask = (1ULL << scale) - 1;
  const uint64_t k = o_mix64(seed ^ 0xA0761D6478BD642FULL);
  uint64_t y = x;
  y = (y + k) & mask;
  y = (y * ((k >> 32) | 1ULL)) & mask;
  y ^= y >> ((scale + 1) / 2);
  y = (y * 0x9E3779B97F4A7C15ULL) & mask;
  y ^= y >> ((scale + 2) / 3);
  y = (y + (k >> 11)) & mask;
  return (uint32_t)y;
}

void o_rmat_scramble_many(int scale, uint64_t seed, uint64_t n, const uint32_t* x, uint32_t* y) {
  for (uint64_t i = 0; i < n; ++i) y[i] = o_rmat_scramble(scale, seed, x[i]);
}

/* flags: 1 = symmetrise, 2 = scramble the labels */
size_t o_rmat_keys(int scale, int ef, uint64_t seed, int flags, uint64_t* keys) {
  const uint64_t ngen = (uint64_t)ef << scale;
  const int sym = flags & 1;
  size_t n = 0;
  for (uint64_t i = 0; i < ngen; ++i) {
    uint32_t u, v;
    rmat_edge(scale, seed, i, &u, &v);
    if (u == v) continue;
    if (flags & 2) { u = o_rmat_scramble(scale, seed, u); v = o_rmat_scramble(scale, seed, v); }
    keys[n++] = ((uint64_t)u << 32) | v;
    if (sym) keys[n++] = ((uint64_t)v << 32) | u;
  }
  o_sort_unique_keys(keys, &n);
  return n;
}

void o_csr_from_keys(uint32_t nV, size_t nE, const uint64_t* keys, uint64_t* off, uint32_t* idx) {
  memset(off, 0, (size_t)(nV + 1) * sizeof(uint64_t));
  for (size_t e = 0; e < nE; ++e) off[(keys[e] >> 32) + 1]++;
  for (uint32_t v = 0; v < nV; ++v) off[v + 1] += off[v];
  for (size_t e = 0; e < nE; ++e) idx[e] = (uint32_t)(keys[e] & 0xffffffffULL);
}

/* GraphInCSR::build (include/Graph/LocalGraph.hpp:66-107): count, scan, fill
 * in input order, so each row keeps the order the edges were read in. */
void o_csr_from_edges(uint32_t nV, size_t nE, const uint32_t* src, const uint32_t* dst,
                      uint64_t* off, uint32_t* idx) {
  memset(off, 0, (size_t)(nV + 1) * sizeof(uint64_t));
  for (size_t e = 0; e < nE; ++e) off[src[e] + 1]++;
  for (uint32_t v = 0; v < nV; ++v) off[v + 1] += off[v];
  uint64_t* cur = (uint64_t*)malloc((size_t)nV * sizeof(uint64_t));
  memcpy(cur, off, (size_t)nV * sizeof(uint64_t));
  for (size_t e = 0; e < nE; ++e) idx[cur[src[e]]++] = dst[e];
  free(cur);
}

/* GraphInCSR::rbuild (include/Graph/LocalGraph.hpp:112-156): reverse CSR whose
 * in-row order is the edge's input position; redgeID keeps that position. */
void o_csc_from_csr(uint32_t nV, const uint64_t* off, const uint32_t* idx,
                    uint64_t* coff, uint32_t* cidx, uint64_t* ceid) {
  const uint64_t nE = off[nV];
  memset(coff, 0, (size_t)(nV + 1) * sizeof(uint64_t));
  for (uint64_t e = 0; e < nE; ++e) coff[idx[e] + 1]++;
  for (uint32_t v = 0; v < nV; ++v) coff[v + 1] += coff[v];
  uint64_t* cur = (uint64_t*)malloc((size_t)nV * sizeof(uint64_t));
  memcpy(cur, coff, (size_t)nV * sizeof(uint64_t));
  for (uint32_t u = 0; u < nV; ++u)
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
      const uint64_t p = cur[idx[e]]++;
      cidx[p] = u;
      if (ceid) ceid[p] = e;
    }
  free(cur);
}

uint32_t o_sssp_weight(uint64_t seed, uint32_t u, uint32_t v) {
  const uint32_t a = u < v ? u : v, b = u < v ? v : u;
  return 1u + (uint32_t)(o_mix64(seed * K0 + (uint64_t)a * K1 + (uint64_t)b * K2 + K3) % 255ULL);
}

void o_sssp_weights_csr(uint64_t seed, uint32_t nV, const uint64_t* off, const uint32_t* idx, uint32_t* w) {
  for (uint32_t u = 0; u < nV; ++u)
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) w[e] = o_sssp_weight(seed, u, idx[e]);
}

double o_lognormal_weight(uint64_t seed, uint32_t u, uint32_t v) {
  const uint64_t h1 = o_mix64(seed * K0 + (uint64_t)u * K1 + (uint64_t)v * K2 + K3);
  const uint64_t h2 = o_mix64(h1 ^ K4);
  const double u1 = ((double)(h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
  const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  return exp(z);
}

/* ------------------------------------------------------------------------ */
/* GRE engine restatement                                                   */
/* ------------------------------------------------------------------------ */
typedef struct { int32_t level; int32_t numPaths; } o_vtx;   /* BC.hpp:12-17 */

typedef struct {
  int prog;
  const o_graph* g;
  void* S; void* C; void* X;
  const void* aux;
  int32_t curr_level;
} o_prog;

/* init(lv) of each shipped vertex program */
static int prog_init(o_prog* p, uint32_t v) {
  switch (p->prog) {
  case O_PAGERANK:          /* PageRank.hpp:38-47 */
  case O_PAGERANK_DELTA: {  /* PageRank.hpp:147-156 */
    double* pr = (double*)p->X; double* ps = (double*)p->S;
    pr[v] = 0.15;
    const int32_t d = p->g->outdeg[v];
    if (d > 0) { ps[v] = pr[v] / (double)d; return 1; }
    return 0;
  }
  case O_BFS: {             /* BFS.hpp:34-42 (np=1: get_vid(lv) == lv) */
    uint64_t* pred = (uint64_t*)p->X; uint64_t* ps = (uint64_t*)p->S;
    pred[v] = v; ps[v] = pred[v];
    return 1;
  }
  case O_CC: {              /* CC.hpp:33-38 */
    uint64_t* cc = (uint64_t*)p->X; uint64_t* cs = (uint64_t*)p->S; uint64_t* cm = (uint64_t*)p->C;
    cc[v] = v; cs[v] = cc[v]; cm[v] = cc[v];
    return 1;
  }
  case O_SSSP: {            /* SSSP.hpp:55-59: states aliases combine and vertex data */
    uint32_t* st = (uint32_t*)p->C; uint32_t* last = (uint32_t*)p->S;
    st[v] = 0; last[v] = st[v];
    return 1;
  }
  case O_BC_FORWARD: {      /* BC.hpp:56-63 */
    o_vtx* s = (o_vtx*)p->X; int32_t* np = (int32_t*)p->S;
    s[v].numPaths = 1; s[v].level = 0; np[v] = s[v].numPaths;
    return 1;
  }
  case O_BC_BACKWARD: {     /* BC.hpp:132-138 */
    const o_vtx* bs = (const o_vtx*)p->aux; double* pw = (double*)p->S; double* dep = (double*)p->X;
    if (bs[v].level == p->curr_level) { pw[v] = (1.0 + dep[v]) / (double)bs[v].numPaths; return 1; }
    return 0;
  }
  case O_COUNTIN: {         /* CountIn.hpp:35-39 */
    int32_t* cn = (int32_t*)p->X; int32_t* cs = (int32_t*)p->S;
    cn[v] = 0; cs[v] = 1;
    return 1;
  }
  }
  return 0;
}

/* scatter(lsrc, ldst, it) followed by sendMessage -> combine(ldst, msg, lock);
 * returns combine()'s result.  epos is the CSR position of the edge. */
static int prog_scatter_combine(o_prog* p, uint32_t u, uint32_t v, uint64_t epos) {
  switch (p->prog) {
  case O_PAGERANK: case O_PAGERANK_DELTA: {   /* PageRank.hpp:52-62 */
    double* cm = (double*)p->C; const double* ps = (const double*)p->S;
    cm[v] += ps[u];
    return 1;
  }
  case O_BFS: {                               /* BFS.hpp:46-66 */
    uint64_t* cm = (uint64_t*)p->C; const uint64_t* ps = (const uint64_t*)p->S;
    if (cm[v] == (uint64_t)-1) { cm[v] = ps[u]; return 1; }
    return 0;
  }
  case O_CC: {                                /* CC.hpp:42-59 */
    uint64_t* cm = (uint64_t*)p->C; const uint64_t* cs = (const uint64_t*)p->S;
    if (cm[v] > cs[u]) { cm[v] = cs[u]; return 1; }
    return 0;
  }
  case O_SSSP: {                              /* SSSP.hpp:63-80 */
    uint32_t* st = (uint32_t*)p->C; const uint32_t* last = (const uint32_t*)p->S;
    const uint32_t* w = (const uint32_t*)p->aux;
    const uint32_t d = last[u] + w[epos];
    if (d < st[v]) { st[v] = d; return 1; }
    return 0;
  }
  case O_BC_FORWARD: {                        /* BC.hpp:66-77; int32 wraps */
    int32_t* sum = (int32_t*)p->C; const int32_t* np = (const int32_t*)p->S;
    sum[v] = (int32_t)((uint32_t)sum[v] + (uint32_t)np[u]);
    return 1;
  }
  case O_BC_BACKWARD: {                       /* BC.hpp:141-152 */
    double* sum = (double*)p->C; const double* pw = (const double*)p->S;
    sum[v] += pw[u];
    return 1;
  }
  case O_COUNTIN: {                           /* CountIn.hpp:42-53 */
    int32_t* cm = (int32_t*)p->C; const int32_t* cs = (const int32_t*)p->S;
    cm[v] += cs[u];
    return 1;
  }
  }
  return 0;
}

static int prog_apply(o_prog* p, uint32_t v) {
  switch (p->prog) {
  case O_PAGERANK: {        /* PageRank.hpp:70-74 */
    double* pr = (double*)p->X; double* cm = (double*)p->C;
    pr[v] = 0.15 + (1.0 - 0.15) * cm[v];
    cm[v] = 0.0;
    return 0;
  }
  case O_PAGERANK_DELTA: {  /* PageRank.hpp:176-189 */
    double* pr = (double*)p->X; double* cm = (double*)p->C; double* ps = (double*)p->S;
    const double delta = (1.0 - 0.15) * cm[v];
    pr[v] += delta;
    cm[v] = 0.0;
    const int32_t d = p->g->outdeg[v];
    if (delta > 1e-5 && d > 0) { ps[v] = delta / (double)d; return 1; }
    return 0;
  }
  case O_BFS: {             /* BFS.hpp:69-75 */
    uint64_t* pred = (uint64_t*)p->X; uint64_t* cm = (uint64_t*)p->C; uint64_t* ps = (uint64_t*)p->S;
    if (pred[v] == (uint64_t)-1) { pred[v] = cm[v]; ps[v] = pred[v]; return 1; }
    return 0;
  }
  case O_CC: {              /* CC.hpp:62-67 */
    uint64_t* cc = (uint64_t*)p->X; uint64_t* cm = (uint64_t*)p->C; uint64_t* cs = (uint64_t*)p->S;
    if (cm[v] >= cc[v]) return 0;
    cc[v] = cm[v]; cs[v] = cc[v];
    return 1;
  }
  case O_SSSP: {            /* SSSP.hpp:83-86 */
    uint32_t* st = (uint32_t*)p->C; uint32_t* last = (uint32_t*)p->S;
    last[v] = st[v];
    return 1;
  }
  case O_BC_FORWARD: {      /* BC.hpp:80-89 */
    o_vtx* s = (o_vtx*)p->X; int32_t* sum = (int32_t*)p->C; int32_t* np = (int32_t*)p->S;
    if (s[v].level < 0) { s[v].level = p->curr_level; s[v].numPaths = sum[v]; np[v] = s[v].numPaths; return 1; }
    return 0;
  }
  case O_BC_BACKWARD: {     /* BC.hpp:155-163 */
    const o_vtx* bs = (const o_vtx*)p->aux; double* sum = (double*)p->C; double* dep = (double*)p->X;
    if (bs[v].level == p->curr_level - 1) dep[v] = sum[v] * (double)bs[v].numPaths;
    return 0;
  }
  case O_COUNTIN: {         /* CountIn.hpp:56-59 */
    int32_t* cn = (int32_t*)p->X; int32_t* cm = (int32_t*)p->C;
    cn[v] = cm[v];
    return 0;
  }
  }
  return 0;
}

/* assert_to_halt: the VertexCode default is true (VertexCode.hpp:50);
 * PageRank overrides it with false (PageRank.hpp:77-79). */
static int prog_halt(const o_prog* p) { return p->prog == O_PAGERANK ? 0 : 1; }

/* reduce(dc) hook: BC Forward increments currLevel (BC.hpp:40-42). */
static void prog_reduce(o_prog* p) { if (p->prog == O_BC_FORWARD) p->curr_level++; }

#define BIT_TEST(b, i) (((b)[(i) >> 6] >> ((i) & 63)) & 1ULL)
#define BIT_SET(b, i) ((b)[(i) >> 6] |= (1ULL << ((i) & 63)))
#define BIT_CLR(b, i) ((b)[(i) >> 6] &= ~(1ULL << ((i) & 63)))

int64_t o_gre_run(int prog, int dir, const o_graph* g, void* S, void* C, void* X,
                  const void* aux, int32_t* level_io, const uint32_t* srcs, int64_t nsrc,
                  int64_t max_ss, int64_t* counts, int64_t counts_cap) {
  const uint32_t nM = g->nV;
  const uint64_t nW = ((uint64_t)nM + 63) / 64 + 1;   /* Bitmap: size/64+1 words */
  uint64_t* curr = (uint64_t*)calloc(nW, 8);
  uint64_t* next = (uint64_t*)calloc(nW, 8);
  uint32_t* queue = (uint32_t*)malloc(((size_t)nM + 1) * sizeof(uint32_t));
  int64_t qn = 0;
  o_prog p = { prog, g, S, C, X, aux, level_io ? *level_io : 1 };
  if (max_ss <= 0) max_ss = 1024;                     /* Engine.hpp:108,148-153 */

  /* init (Engine.hpp:252-257 master pre-insert, :437-471 workers) */
  if (!srcs) {
    for (uint32_t v = 0; v < nM; ++v)
      if (prog_init(&p, v)) { BIT_SET(curr, v); queue[qn++] = v; }
  } else {
    for (int64_t i = 0; i < nsrc; ++i) {
      const uint32_t s = srcs[i];
      if (!BIT_TEST(curr, s)) { BIT_SET(curr, s); queue[qn++] = s; }
    }
    for (int64_t i = 0; i < nsrc; ++i) {
      const uint32_t s = srcs[i];
      if (prog_init(&p, s) && !BIT_TEST(curr, s)) { BIT_SET(curr, s); queue[qn++] = s; }
    }
  }

  int64_t ss;
  for (ss = 0; ss < max_ss; ++ss) {
    memset(next, 0, nW * 8);                          /* nextSet.clear() :322 */
    /* compute (Engine.hpp:592-625): batches of scheduleGrain = 8 */
    for (int64_t b = 0; b < qn; b += 8) {
      const int64_t be = b + 8 < qn ? b + 8 : qn;
      if (dir == O_DIR_OUT || dir == O_DIR_INOUT)
        for (int64_t i = b; i < be; ++i) {
          const uint32_t u = queue[i];
          for (uint64_t e = g->off[u]; e < g->off[u + 1]; ++e)
            if (prog_scatter_combine(&p, u, g->idx[e], e)) BIT_SET(next, g->idx[e]);
        }
      if (dir == O_DIR_IN || dir == O_DIR_INOUT)
        for (int64_t i = b; i < be; ++i) {
          const uint32_t u = queue[i];
          for (uint64_t e = g->coff[u]; e < g->coff[u + 1]; ++e)
            if (prog_scatter_combine(&p, u, g->cidx[e], g->ceid ? g->ceid[e] : e)) BIT_SET(next, g->cidx[e]);
        }
      if (prog_halt(&p))
        for (int64_t i = b; i < be; ++i) BIT_CLR(curr, queue[i]);
    }
    qn = 0;                                           /* currSet.clear_set() :377 */
    /* apply (Engine.hpp:757-775) */
    for (uint32_t v = 0; v < nM; ++v)
      if (BIT_TEST(next, v) && prog_apply(&p, v)) BIT_SET(curr, v);
    /* rebuild currSet vector from the whole master bitmap (:777-798) */
    for (uint32_t v = 0; v < nM; ++v)
      if (BIT_TEST(curr, v)) queue[qn++] = v;
    prog_reduce(&p);                                  /* vertexCode.reduce(dc) :390 */
    if (counts && ss < counts_cap) counts[ss] = qn;
    if (qn == 0) break;                               /* :399-400 */
  }
  if (level_io) *level_io = p.curr_level;
  free(curr); free(next); free(queue);
  return ss;
}

/* ------------------------------------------------------------------------ */
/* Textbook references                                                      */
/* ------------------------------------------------------------------------ */
void o_bfs_levels(uint32_t nV, const uint64_t* off, const uint32_t* idx, uint32_t src, int32_t* level) {
  uint32_t* q = (uint32_t*)malloc((size_t)nV * sizeof(uint32_t));
  for (uint32_t v = 0; v < nV; ++v) level[v] = -1;
  size_t h = 0, t = 0;
  level[src] = 0; q[t++] = src;
  while (h < t) {
    const uint32_t u = q[h++];
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
      const uint32_t v = idx[e];
      if (level[v] < 0) { level[v] = level[u] + 1; q[t++] = v; }
    }
  }
  free(q);
}

static uint32_t uf_find(uint32_t* par, uint32_t x) {
  while (par[x] != x) { par[x] = par[par[x]]; x = par[x]; }
  return x;
}

void o_cc_unionfind(uint32_t nV, const uint64_t* off, const uint32_t* idx, uint32_t* label) {
  uint32_t* par = (uint32_t*)malloc((size_t)nV * sizeof(uint32_t));
  for (uint32_t v = 0; v < nV; ++v) par[v] = v;
  for (uint32_t u = 0; u < nV; ++u)
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
      uint32_t a = uf_find(par, u), b = uf_find(par, idx[e]);
      if (a == b) continue;
      if (a < b) par[b] = a; else par[a] = b;      /* root = smallest id */
    }
  for (uint32_t v = 0; v < nV; ++v) label[v] = uf_find(par, v);
  free(par);
}

/* binary heap keyed on (dist, vertex) */
typedef struct { uint64_t* h; size_t n; } o_heap;
static void heap_push(o_heap* hp, uint64_t k) {
  size_t i = hp->n++;
  hp->h[i] = k;
  while (i > 0) {
    size_t pa = (i - 1) / 2;
    if (hp->h[pa] <= hp->h[i]) break;
    uint64_t t = hp->h[pa]; hp->h[pa] = hp->h[i]; hp->h[i] = t; i = pa;
  }
}
static uint64_t heap_pop(o_heap* hp) {
  uint64_t top = hp->h[0];
  hp->h[0] = hp->h[--hp->n];
  size_t i = 0;
  for (;;) {
    size_t l = 2 * i + 1, r = l + 1, m = i;
    if (l < hp->n && hp->h[l] < hp->h[m]) m = l;
    if (r < hp->n && hp->h[r] < hp->h[m]) m = r;
    if (m == i) break;
    uint64_t t = hp->h[m]; hp->h[m] = hp->h[i]; hp->h[i] = t; i = m;
  }
  return top;
}

void o_dijkstra(uint32_t nV, const uint64_t* off, const uint32_t* idx, const uint32_t* w,
                uint32_t src, uint32_t* dist) {
  o_heap hp;
  hp.h = (uint64_t*)malloc((off[nV] + 1) * sizeof(uint64_t));
  hp.n = 0;
  for (uint32_t v = 0; v < nV; ++v) dist[v] = 0xffffffffu;
  dist[src] = 0;
  heap_push(&hp, (uint64_t)src);
  while (hp.n) {
    const uint64_t k = heap_pop(&hp);
    const uint32_t u = (uint32_t)(k & 0xffffffffULL), du = (uint32_t)(k >> 32);
    if (du != dist[u]) continue;
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
      const uint32_t v = idx[e];
      const uint32_t nd = du + w[e];
      if (nd < dist[v]) { dist[v] = nd; heap_push(&hp, ((uint64_t)nd << 32) | v); }
    }
  }
  free(hp.h);
}

void o_brandes_one(uint32_t nV, const uint64_t* off, const uint32_t* idx, uint32_t src,
                   double* sigma, double* delta, int32_t* level) {
  uint32_t* order = (uint32_t*)malloc((size_t)nV * sizeof(uint32_t));
  for (uint32_t v = 0; v < nV; ++v) { level[v] = -1; sigma[v] = 0.0; delta[v] = 0.0; }
  size_t h = 0, t = 0;
  level[src] = 0; sigma[src] = 1.0; order[t++] = src;
  while (h < t) {
    const uint32_t u = order[h++];
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
      const uint32_t v = idx[e];
      if (level[v] < 0) { level[v] = level[u] + 1; order[t++] = v; }
      if (level[v] == level[u] + 1) sigma[v] += sigma[u];
    }
  }
  for (size_t i = t; i-- > 0;) {
    const uint32_t u = order[i];
    double acc = 0.0;
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
      const uint32_t v = idx[e];
      if (level[v] == level[u] + 1) acc += (sigma[u] / sigma[v]) * (1.0 + delta[v]);
    }
    delta[u] = acc;
  }
  free(order);
}

void o_pagerank(uint32_t nV, const uint64_t* coff, const uint32_t* cidx, const int32_t* outdeg,
                double damping, int iters, double* x) {
  double* contrib = (double*)malloc((size_t)nV * sizeof(double));
  const double invn = 1.0 / (double)nV;
  for (uint32_t v = 0; v < nV; ++v) x[v] = invn;
  for (int it = 0; it < iters; ++it) {
    double dang = 0.0;
    for (uint32_t v = 0; v < nV; ++v) {
      if (outdeg[v] > 0) contrib[v] = x[v] / (double)outdeg[v];
      else { contrib[v] = 0.0; dang += x[v]; }
    }
    const double base = (1.0 - damping) * invn + damping * dang * invn;
    for (uint32_t v = 0; v < nV; ++v) {
      double s = 0.0;
      for (uint64_t e = coff[v]; e < coff[v + 1]; ++e) s += contrib[cidx[e]];
      x[v] = base + damping * s;
    }
  }
  free(contrib);
}

int o_ppr(uint32_t nV, const uint64_t* coff, const uint32_t* cidx, const double* wc,
          const int32_t* outdeg, const double* p, double damping, double tol, int maxit, double* x) {
  double* y = (double*)malloc((size_t)nV * sizeof(double));
  for (uint32_t v = 0; v < nV; ++v) x[v] = p[v];
  int it;
  for (it = 0; it < maxit; ++it) {
    double dang = 0.0;
#pragma omp parallel for reduction(+:dang) schedule(static)
    for (int64_t v = 0; v < (int64_t)nV; ++v) if (outdeg[v] == 0) dang += x[v];
    double diff = 0.0;
    /* rows are independent; each row sums in CSC order whatever the thread count */
#pragma omp parallel for reduction(+:diff) schedule(dynamic, 4096)
    for (int64_t v = 0; v < (int64_t)nV; ++v) {
      double s = 0.0;
      for (uint64_t e = coff[v]; e < coff[v + 1]; ++e) s += wc[e] * x[cidx[e]];
      y[v] = damping * (s + dang * p[v]) + (1.0 - damping) * p[v];
      diff += fabs(y[v] - x[v]);
    }
    memcpy(x, y, (size_t)nV * sizeof(double));
    if (diff < tol) { ++it; break; }
  }
  free(y);
  return it;
}

int o_ppr_adjoint(uint32_t nV, const uint64_t* off, const uint32_t* idx, const double* wr,
                  const int32_t* outdeg, const double* p, const double* c, double damping,
                  double tol, int maxit, double* lam) {
  double* y = (double*)malloc((size_t)nV * sizeof(double));
  for (uint32_t v = 0; v < nV; ++v) lam[v] = c[v];
  int it;
  for (it = 0; it < maxit; ++it) {
    double plam = 0.0;
#pragma omp parallel for reduction(+:plam) schedule(static)
    for (int64_t v = 0; v < (int64_t)nV; ++v) plam += p[v] * lam[v];
    double diff = 0.0;
#pragma omp parallel for reduction(+:diff) schedule(dynamic, 4096)
    for (int64_t u = 0; u < (int64_t)nV; ++u) {
      double s = 0.0;
      for (uint64_t e = off[u]; e < off[u + 1]; ++e) s += wr[e] * lam[idx[e]];
      if (outdeg[u] == 0) s += plam;
      y[u] = c[u] + damping * s;
      diff += fabs(y[u] - lam[u]);
    }
    memcpy(lam, y, (size_t)nV * sizeof(double));
    if (diff < tol) { ++it; break; }
  }
  free(y);
  return it;
}

void o_ppr_edge_grad(uint32_t nV, const uint64_t* off, const uint32_t* idx, const double* wr,
                     const double* rowsum, const double* x, const double* lam, double damping,
                     double* gW, double* graw) {
  for (uint32_t u = 0; u < nV; ++u) {
    double wl = 0.0;
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) wl += wr[e] * lam[idx[e]];
    for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
      const double g = damping * x[u] * lam[idx[e]];
      gW[e] = g;
      graw[e] = rowsum[u] > 0.0 ? damping * x[u] * (lam[idx[e]] - wl) / rowsum[u] : 0.0;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* GRE-style multi-threaded push superstep (CPU baseline timing)            */
/* ------------------------------------------------------------------------ */
static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

double o_gre_pagerank_mt(int nthreads, uint32_t nV, const uint64_t* off, const uint32_t* idx,
                         const int32_t* outdeg, double damping, int iters, double* x) {
  double* ps = (double*)malloc((size_t)nV * sizeof(double));
  double* cm = (double*)calloc((size_t)nV, sizeof(double));
  volatile int32_t* locks = (volatile int32_t*)calloc(8192, sizeof(int32_t));
  const double invn = 1.0 / (double)nV;
  for (uint32_t v = 0; v < nV; ++v) x[v] = invn;
  if (nthreads < 1) nthreads = 1;
  const double t0 = now_s();
  for (int it = 0; it < iters; ++it) {
    double dang = 0.0;
    /* scatter data refresh + dangling mass (the vertex program's reduce) */
#pragma omp parallel for num_threads(nthreads) reduction(+:dang) schedule(static)
    for (int64_t v = 0; v < (int64_t)nV; ++v) {
      if (outdeg[v] > 0) ps[v] = x[v] / (double)outdeg[v];
      else { ps[v] = 0.0; dang += x[v]; }
    }
    /* compute phase: dynamic grain-8 fetch, per-edge locked combine */
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 8)
    for (int64_t u = 0; u < (int64_t)nV; ++u) {
      if (outdeg[u] == 0) continue;
      const double m = ps[u];
      for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
        const uint32_t v = idx[e];
        volatile int32_t* lk = &locks[v & 8191];
        while (!__sync_bool_compare_and_swap(lk, 0, 1)) { while (*lk == 1) {} }
        cm[v] += m;
        __sync_synchronize();
        *lk = 0;
      }
    }
    /* apply phase */
    const double base = (1.0 - damping) * invn + damping * dang * invn;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t v = 0; v < (int64_t)nV; ++v) { x[v] = base + damping * cm[v]; cm[v] = 0.0; }
  }
  const double t1 = now_s();
  free(ps); free(cm); free((void*)locks);
  return t1 - t0;
}

/* ------------------------------------------------------------------------ */
/* Vertex-cut ingress at np processes (include/Graph/GraphIngress.hpp)      */
/* ------------------------------------------------------------------------ */
/* LDGEdgeDecision::getPartID (GraphIngress.hpp:1144-1170), one decision
 * state per reading process; process r reads files r, r+np, ... in order
 * (loadEdgesFromFiles :511-600), capacity nE/np/np (:560).  Returns -1 when
 * the capacity is 0 (GRE divides by it). */
/* One decision: the part with the highest balance-weighted locality score,
 * ties (within 1e-4) broken by (src + dst) % #tied (GraphIngress.hpp:1144-1170). */
static int ldg_pick(int np, uint64_t capacity, const uint64_t* cnt, uint64_t sb, uint64_t db, uint64_t s,
                    uint64_t d) {
  double sc[64];
  int cand[64];
  for (int i = 0; i < np; ++i) {
    const double bal = fabs(1.0 - (double)cnt[i] / (double)capacity);
    /* (srcMap.test(i) ? 1.0 : (src%np==i ? 0.5 : 0.0) + (dstMap.test(i) ? 1.0 : (dst%np==i ? 0.5 : 0.0))) */
    double term;
    if ((sb >> i) & 1ULL) {
      term = 1.0;
    } else {
      term = ((int)(s % (uint64_t)np) == i) ? 0.5 : 0.0;
      if ((db >> i) & 1ULL) term += 1.0;
      else if ((int)(d % (uint64_t)np) == i) term += 0.5;
    }
    sc[i] = bal * term;
  }
  double mx = sc[0];                             /* Util::getMax: first, then > */
  for (int i = 0; i < np; ++i) if (sc[i] > mx) mx = sc[i];
  int n = 0;
  for (int i = 0; i < np; ++i) if (fabs(mx - sc[i]) < 1e-4) cand[n++] = i;
  return cand[(s + d) % (uint64_t)n];
}

int o_ldg_assign(uint32_t nV, uint64_t nE_meta, int np, int nfiles, const uint64_t* file_off,
                 const uint32_t* src, const uint32_t* dst, int32_t* part) {
  const uint64_t capacity = nE_meta / (uint64_t)np / (uint64_t)np;
  if (capacity == 0) return -1;
  uint64_t* sbits = (uint64_t*)malloc((size_t)nV * 8);
  uint64_t* dbits = (uint64_t*)malloc((size_t)nV * 8);
  uint64_t cnt[64];
  for (int reader = 0; reader < np; ++reader) {
    memset(sbits, 0, (size_t)nV * 8);
    memset(dbits, 0, (size_t)nV * 8);
    memset(cnt, 0, sizeof(cnt));
    for (int f = reader; f < nfiles; f += np) {
      for (uint64_t e = file_off[f]; e < file_off[f + 1]; ++e) {
        const uint64_t s = src[e], d = dst[e];
        const int pid = ldg_pick(np, capacity, cnt, sbits[s], dbits[d], s, d);
        cnt[pid]++;
        sbits[s] |= 1ULL << pid;
        dbits[d] |= 1ULL << pid;
        part[e] = pid;
      }
    }
  }
  free(sbits); free(dbits);
  return 0;
}

/* Per part p: masters, scatters, combiners, bi-agents, agents, local edges,
 * aligned masters (GraphIngress.hpp:111-153, 229-235; getPartID = vid % np). */
void o_vertex_cut_stats(uint32_t nV, int np, int dir, uint64_t nE, const uint32_t* src, const uint32_t* dst,
                        const int32_t* part, int64_t* stats) {
  uint8_t* sc = (uint8_t*)malloc(nV ? nV : 1);
  uint8_t* cb = (uint8_t*)malloc(nV ? nV : 1);
  for (int p = 0; p < np; ++p) {
    memset(sc, 0, nV); memset(cb, 0, nV);
    int64_t ne = 0, ns = 0, nc = 0, nb = 0;
    for (uint64_t e = 0; e < nE; ++e) {
      if (part[e] != p) continue;
      ne++;
      if (src[e] % (uint32_t)np != (uint32_t)p) sc[src[e]] = 1;
      if (dst[e] % (uint32_t)np != (uint32_t)p) cb[dst[e]] = 1;
    }
    for (uint32_t v = 0; v < nV; ++v) { ns += sc[v]; nc += cb[v]; nb += sc[v] & cb[v]; }
    const int64_t nm = nV / (uint32_t)np + ((uint32_t)p < nV % (uint32_t)np);
    const int64_t ag = nc + ns - nb;
    int64_t* o = stats + 7 * p;
    o[0] = nm;
    o[1] = dir == O_DIR_OUT ? ns : ag;
    o[2] = dir == O_DIR_OUT ? nc : ag;
    o[3] = nb;
    o[4] = ag;
    o[5] = ne;
    o[6] = (nm + 63) / 64 * 64;
  }
  free(sc); free(cb);
}

/* ------------------------------------------------------------------------ */
/* GRE's PageRank superstep at np processes on one host (CPU baseline)      */
/* ------------------------------------------------------------------------ */
/* Ingress: edge file i holds the CSR rows u with u % np == i, in order (the
 * layout GRE's out-degree counting assumes, GraphIngress.hpp:591,596);
 * reader r takes file r and places each edge by LDG (ldg_pick), capacity
 * E/np/np.  Part p keeps its masters (v % np == p, ascending) then its agents
 * (ascending), a local CSR of its edges, and the send lists of the two syncs.
 * A superstep follows Engine.hpp:427-700: masters refresh their scatter data,
 * scatter sync (each master's value to the parts holding it as a scatter
 * agent), local push over the part's edges (dynamic grain-8 fetch, striped
 * locks), combiner sync (each combiner agent's sum to its master, then reset),
 * apply on the masters.  Processes are thread groups with private arrays; a
 * message is (vid, value) packed into the sender's buffer and unpacked by the
 * receiver through its own vid -> local id map, with a barrier between
 * phases, as MPI moves them through shared memory on one host.  The apply is
 * the textbook one with the dangling mass (as o_gre_pagerank_mt). */
typedef struct { uint64_t vid; double val; } o_msg;
typedef struct {
  uint32_t nl, nm;
  uint32_t* l2g;
  uint32_t* g2l;          /* [nV], UINT32_MAX: not local */
  uint64_t* off;          /* [nl+1] local out-edges by local source */
  uint32_t* adj;
  uint64_t ss_off[65];    /* scatter sends to part q: ss_lv[ss_off[q] .. ss_off[q+1]) */
  uint32_t* ss_lv;
  uint64_t cs_off[65];    /* combiner sends to master part q */
  uint32_t* cs_lv;
  o_msg* sbuf;            /* [ss_off[np]] */
  o_msg* cbuf;            /* [cs_off[np]] */
  double *ps, *acc, *x;
  int32_t* deg;           /* masters' out-degree */
  volatile int32_t* locks;
  volatile int64_t next;
} o_part;

typedef struct {
  int np, W, iters;
  uint32_t nV;
  double damping;
  o_part* parts;
  double* dang;           /* [np*W] per-thread dangling partials */
  pthread_barrier_t bar;
  double phase[5];        /* refresh, scatter sync, compute, combiner sync, apply (thread 0 of part 0) */
} o_np_run;

typedef struct { o_np_run* r; int p, t; } o_np_arg;

static void lock_add(volatile int32_t* locks, uint32_t k, double* a, double v) {
  volatile int32_t* lk = &locks[k & 8191];
  while (!__sync_bool_compare_and_swap(lk, 0, 1)) { while (*lk == 1) {} }
  *a += v;
  __sync_synchronize();
  *lk = 0;
}

static void* np_worker(void* argp) {
  o_np_arg* a = (o_np_arg*)argp;
  o_np_run* r = a->r;
  const int np = r->np, W = r->W, p = a->p, t = a->t;
  o_part* P = &r->parts[p];
  const int lead = p == 0 && t == 0;
  const double invn = 1.0 / (double)r->nV;
#define SLICE(n, lo, hi) const uint64_t lo = (uint64_t)(n) * t / W, hi = (uint64_t)(n) * (t + 1) / W
  for (int it = 0; it < r->iters; ++it) {
    double t0 = lead ? now_s() : 0.0;
    /* masters: scatter data and dangling mass */
    {
      SLICE(P->nm, lo, hi);
      double dg = 0.0;
      for (uint64_t i = lo; i < hi; ++i) {
        if (P->deg[i] > 0) P->ps[i] = P->x[i] / (double)P->deg[i];
        else { P->ps[i] = 0.0; dg += P->x[i]; }
      }
      r->dang[p * W + t] = dg;
      if (t == 0) P->next = 0;
    }
    pthread_barrier_wait(&r->bar);
    double t1 = lead ? now_s() : 0.0;
    /* scatter sync: pack, then unpack the messages addressed to this part */
    {
      SLICE(P->ss_off[np], lo, hi);
      for (uint64_t i = lo; i < hi; ++i) {
        const uint32_t lv = P->ss_lv[i];
        P->sbuf[i].vid = P->l2g[lv];
        P->sbuf[i].val = P->ps[lv];
      }
    }
    pthread_barrier_wait(&r->bar);
    for (int q = 0; q < np; ++q) {
      if (q == p) continue;
      const o_part* Q = &r->parts[q];
      const uint64_t b = Q->ss_off[p], n = Q->ss_off[p + 1] - b;
      SLICE(n, lo, hi);
      for (uint64_t i = lo; i < hi; ++i) {
        const o_msg m = Q->sbuf[b + i];
        P->ps[P->g2l[m.vid]] = m.val;
      }
    }
    pthread_barrier_wait(&r->bar);
    double t2 = lead ? now_s() : 0.0;
    /* local push over the part's edges */
    for (;;) {
      const int64_t u0 = __sync_fetch_and_add(&P->next, 8);
      if (u0 >= (int64_t)P->nl) break;
      const int64_t u1 = u0 + 8 < (int64_t)P->nl ? u0 + 8 : (int64_t)P->nl;
      for (int64_t u = u0; u < u1; ++u) {
        if (P->off[u] == P->off[u + 1]) continue;
        const double m = P->ps[u];
        for (uint64_t e = P->off[u]; e < P->off[u + 1]; ++e) {
          const uint32_t v = P->adj[e];
          lock_add(P->locks, v, &P->acc[v], m);
        }
      }
    }
    pthread_barrier_wait(&r->bar);
    double t3 = lead ? now_s() : 0.0;
    /* combiner sync: pack and reset the agents' sums, then add the received ones */
    {
      SLICE(P->cs_off[np], lo, hi);
      for (uint64_t i = lo; i < hi; ++i) {
        const uint32_t lv = P->cs_lv[i];
        P->cbuf[i].vid = P->l2g[lv];
        P->cbuf[i].val = P->acc[lv];
        P->acc[lv] = 0.0;
      }
    }
    pthread_barrier_wait(&r->bar);
    for (int q = 0; q < np; ++q) {
      if (q == p) continue;
      const o_part* Q = &r->parts[q];
      const uint64_t b = Q->cs_off[p], n = Q->cs_off[p + 1] - b;
      SLICE(n, lo, hi);
      for (uint64_t i = lo; i < hi; ++i) {
        const o_msg m = Q->cbuf[b + i];
        const uint32_t lv = (uint32_t)(m.vid / (uint64_t)np);
        lock_add(P->locks, lv, &P->acc[lv], m.val);
      }
    }
    pthread_barrier_wait(&r->bar);
    double t4 = lead ? now_s() : 0.0;
    /* apply on the masters (the dangling mass summed in a fixed order) */
    {
      double dang = 0.0;
      for (int k = 0; k < np * W; ++k) dang += r->dang[k];
      const double base = (1.0 - r->damping) * invn + r->damping * dang * invn;
      SLICE(P->nm, lo, hi);
      for (uint64_t i = lo; i < hi; ++i) {
        P->x[i] = base + r->damping * P->acc[i];
        P->acc[i] = 0.0;
      }
    }
    pthread_barrier_wait(&r->bar);
    if (lead) {
      const double t5 = now_s();
      r->phase[0] += t1 - t0; r->phase[1] += t2 - t1; r->phase[2] += t3 - t2;
      r->phase[3] += t4 - t3; r->phase[4] += t5 - t4;
    }
  }
#undef SLICE
  return NULL;
}

/* One ingress, then `iters` supersteps from x = 1/V once per worker count
 * workers[k] (k < nw), secs[k] their seconds; x holds the last run's ranks.
 * info[0..4]: the last run's phase seconds (refresh, scatter sync, compute,
 * combiner sync, apply); info[5]: ingress seconds; info[6], info[7]: scatter
 * and combiner messages per superstep; info[8 + p]: edges of part p.
 * Returns 0, or -1 (np out of [1, 64], a worker count below 1, capacity 0,
 * or no memory). */
int o_gre_pagerank_np(int np, int nw, const int* workers, uint32_t nV, const uint64_t* off, const uint32_t* idx,
                      const int32_t* outdeg, double damping, int iters, double* x, double* secs, double* info) {
  const uint64_t nE = off[nV];
  if (np < 1 || np > 64 || nw < 1) return -1;
  for (int k = 0; k < nw; ++k) if (workers[k] < 1) return -1;
  const uint64_t capacity = nE / (uint64_t)np / (uint64_t)np;
  if (capacity == 0) return -1;
  const double ti0 = now_s();
  int32_t* part = (int32_t*)malloc((nE ? nE : 1) * sizeof(int32_t));
  uint8_t* smark = (uint8_t*)calloc((size_t)np * nV + 1, 1);   /* v is a source of part p's edges */
  uint8_t* dmark = (uint8_t*)calloc((size_t)np * nV + 1, 1);   /* v is a target of part p's edges */
  o_part* parts = (o_part*)calloc((size_t)np, sizeof(o_part));
  if (!part || !smark || !dmark || !parts) { free(part); free(smark); free(dmark); free(parts); return -1; }
  /* reader r: file r = rows u = r, r + np, ... */
#pragma omp parallel for num_threads(np) schedule(static, 1)
  for (int rd = 0; rd < np; ++rd) {
    uint64_t* sb = (uint64_t*)calloc(nV ? nV : 1, 8);
    uint64_t* db = (uint64_t*)calloc(nV ? nV : 1, 8);
    uint64_t cnt[64] = {0};
    for (uint64_t u = (uint64_t)rd; u < nV; u += (uint64_t)np)
      for (uint64_t e = off[u]; e < off[u + 1]; ++e) {
        const uint64_t d = idx[e];
        const int pid = ldg_pick(np, capacity, cnt, sb[u], db[d], u, d);
        cnt[pid]++;
        sb[u] |= 1ULL << pid;
        db[d] |= 1ULL << pid;
        part[e] = pid;
      }
    free(sb); free(db);
  }
#pragma omp parallel for num_threads(np) schedule(static, 1)
  for (int p = 0; p < np; ++p) {
    uint8_t* sm = smark + (size_t)p * nV;
    uint8_t* dm = dmark + (size_t)p * nV;
    uint64_t ne = 0;
    for (uint64_t u = 0; u < nV; ++u)
      for (uint64_t e = off[u]; e < off[u + 1]; ++e)
        if (part[e] == p) { sm[u] = 1; dm[idx[e]] = 1; ne++; }
    o_part* P = &parts[p];
    P->nm = nV / (uint32_t)np + ((uint32_t)p < nV % (uint32_t)np);
    uint32_t nl = P->nm;
    for (uint32_t v = 0; v < nV; ++v)
      if ((int)(v % (uint32_t)np) != p && (sm[v] || dm[v])) nl++;
    P->nl = nl;
    P->l2g = (uint32_t*)malloc(((size_t)nl + 1) * 4);
    P->g2l = (uint32_t*)malloc(((size_t)nV + 1) * 4);
    memset(P->g2l, 0xff, ((size_t)nV + 1) * 4);
    uint32_t k = 0;
    for (uint64_t v = (uint64_t)p; v < nV; v += (uint64_t)np) { P->l2g[k] = (uint32_t)v; P->g2l[v] = k; k++; }
    for (uint32_t v = 0; v < nV; ++v)
      if ((int)(v % (uint32_t)np) != p && (sm[v] || dm[v])) { P->l2g[k] = v; P->g2l[v] = k; k++; }
    P->off = (uint64_t*)calloc((size_t)nl + 1, 8);
    P->adj = (uint32_t*)malloc((ne ? ne : 1) * 4);
    for (uint64_t u = 0; u < nV; ++u)
      for (uint64_t e = off[u]; e < off[u + 1]; ++e)
        if (part[e] == p) P->off[P->g2l[u] + 1]++;
    for (uint32_t i = 0; i < nl; ++i) P->off[i + 1] += P->off[i];
    uint64_t* fill = (uint64_t*)malloc(((size_t)nl + 1) * 8);
    memcpy(fill, P->off, ((size_t)nl + 1) * 8);
    for (uint64_t u = 0; u < nV; ++u)
      for (uint64_t e = off[u]; e < off[u + 1]; ++e)
        if (part[e] == p) P->adj[fill[P->g2l[u]]++] = P->g2l[idx[e]];
    free(fill);
    P->ps = (double*)calloc((size_t)nl + 1, 8);
    P->acc = (double*)calloc((size_t)nl + 1, 8);
    P->x = (double*)malloc(((size_t)P->nm + 1) * 8);
    P->deg = (int32_t*)malloc(((size_t)P->nm + 1) * 4);
    for (uint32_t i = 0; i < P->nm; ++i) { P->x[i] = 1.0 / (double)nV; P->deg[i] = outdeg[P->l2g[i]]; }
    P->locks = (volatile int32_t*)calloc(8192, sizeof(int32_t));
    if (info) info[8 + p] = (double)ne;
  }
  /* send lists (every part's marks are known now) */
  double nsm = 0.0, ncm = 0.0;
  for (int p = 0; p < np; ++p) {
    o_part* P = &parts[p];
    uint64_t n = 0;
    P->ss_off[0] = 0;
    for (int q = 0; q < np; ++q) {
      if (q != p)
        for (uint32_t i = 0; i < P->nm; ++i) n += smark[(size_t)q * nV + P->l2g[i]];
      P->ss_off[q + 1] = n;
    }
    P->ss_lv = (uint32_t*)malloc((n ? n : 1) * 4);
    n = 0;
    for (int q = 0; q < np; ++q)
      if (q != p)
        for (uint32_t i = 0; i < P->nm; ++i)
          if (smark[(size_t)q * nV + P->l2g[i]]) P->ss_lv[n++] = i;
    P->sbuf = (o_msg*)malloc((n ? n : 1) * sizeof(o_msg));
    nsm += (double)n;
    /* combiner agents, grouped by master part, ascending within */
    uint64_t c = 0;
    P->cs_off[0] = 0;
    const uint8_t* dm = dmark + (size_t)p * nV;
    for (int q = 0; q < np; ++q) {
      if (q != p)
        for (uint64_t v = (uint64_t)q; v < nV; v += (uint64_t)np) c += dm[v];
      P->cs_off[q + 1] = c;
    }
    P->cs_lv = (uint32_t*)malloc((c ? c : 1) * 4);
    c = 0;
    for (int q = 0; q < np; ++q)
      if (q != p)
        for (uint64_t v = (uint64_t)q; v < nV; v += (uint64_t)np)
          if (dm[v]) P->cs_lv[c++] = P->g2l[v];
    P->cbuf = (o_msg*)malloc((c ? c : 1) * sizeof(o_msg));
    ncm += (double)c;
  }
  free(part); free(smark); free(dmark);
  const double ti1 = now_s();

  o_np_run r;
  for (int k = 0; k < nw; ++k) {
    const int W = workers[k];
    for (int p = 0; p < np; ++p) {
      o_part* P = &parts[p];
      for (uint32_t i = 0; i < P->nm; ++i) P->x[i] = 1.0 / (double)nV;
      memset(P->acc, 0, ((size_t)P->nl + 1) * 8);
    }
    memset(&r, 0, sizeof(r));
    r.np = np; r.W = W; r.iters = iters; r.nV = nV; r.damping = damping; r.parts = parts;
    r.dang = (double*)calloc((size_t)np * W, 8);
    pthread_barrier_init(&r.bar, NULL, (unsigned)(np * W));
    pthread_t* th = (pthread_t*)malloc((size_t)np * W * sizeof(pthread_t));
    o_np_arg* args = (o_np_arg*)malloc((size_t)np * W * sizeof(o_np_arg));
    const double t0 = now_s();
    for (int j = 0; j < np * W; ++j) {
      args[j].r = &r; args[j].p = j / W; args[j].t = j % W;
      pthread_create(&th[j], NULL, np_worker, &args[j]);
    }
    for (int j = 0; j < np * W; ++j) pthread_join(th[j], NULL);
    secs[k] = now_s() - t0;
    pthread_barrier_destroy(&r.bar);
    free(r.dang); free(th); free(args);
  }
  for (int p = 0; p < np; ++p) {
    o_part* P = &parts[p];
    for (uint32_t i = 0; i < P->nm; ++i) x[P->l2g[i]] = P->x[i];
    free(P->l2g); free(P->g2l); free(P->off); free(P->adj); free(P->ss_lv); free(P->cs_lv);
    free(P->sbuf); free(P->cbuf); free(P->ps); free(P->acc); free(P->x); free(P->deg); free((void*)P->locks);
  }
  if (info) {
    for (int k = 0; k < 5; ++k) info[k] = r.phase[k];
    info[5] = ti1 - ti0;
    info[6] = nsm;
    info[7] = ncm;
  }
  free(parts);
  return 0;
}
