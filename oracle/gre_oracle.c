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
 *   PageRank_delta near its 1e-5 activation threshold: parity unpinned.
 */
#define _POSIX_C_SOURCE 200809L
#include "gre_oracle.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <time.h>
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

size_t o_rmat_keys(int scale, int ef, uint64_t seed, int sym, uint64_t* keys) {
  const uint64_t ngen = (uint64_t)ef << scale;
  size_t n = 0;
  for (uint64_t i = 0; i < ngen; ++i) {
    uint32_t u, v;
    rmat_edge(scale, seed, i, &u, &v);
    if (u == v) continue;
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
int o_ldg_assign(uint32_t nV, uint64_t nE_meta, int np, int nfiles, const uint64_t* file_off,
                 const uint32_t* src, const uint32_t* dst, int32_t* part) {
  const uint64_t capacity = nE_meta / (uint64_t)np / (uint64_t)np;
  if (capacity == 0) return -1;
  uint64_t* sbits = (uint64_t*)malloc((size_t)nV * 8);
  uint64_t* dbits = (uint64_t*)malloc((size_t)nV * 8);
  uint64_t cnt[64];
  double sc[64];
  int cand[64];
  for (int reader = 0; reader < np; ++reader) {
    memset(sbits, 0, (size_t)nV * 8);
    memset(dbits, 0, (size_t)nV * 8);
    memset(cnt, 0, sizeof(cnt));
    for (int f = reader; f < nfiles; f += np) {
      for (uint64_t e = file_off[f]; e < file_off[f + 1]; ++e) {
        const uint64_t s = src[e], d = dst[e];
        for (int i = 0; i < np; ++i) {
          const double bal = fabs(1.0 - (double)cnt[i] / (double)capacity);
          /* (srcMap.test(i) ? 1.0 : (src%np==i ? 0.5 : 0.0) + (dstMap.test(i) ? 1.0 : (dst%np==i ? 0.5 : 0.0))) */
          double term;
          if ((sbits[s] >> i) & 1ULL) {
            term = 1.0;
          } else {
            term = ((int)(s % (uint64_t)np) == i) ? 0.5 : 0.0;
            if ((dbits[d] >> i) & 1ULL) term += 1.0;
            else if ((int)(d % (uint64_t)np) == i) term += 0.5;
          }
          sc[i] = bal * term;
        }
        double mx = sc[0];                             /* Util::getMax: first, then > */
        for (int i = 0; i < np; ++i) if (sc[i] > mx) mx = sc[i];
        int n = 0;
        for (int i = 0; i < np; ++i) if (fabs(mx - sc[i]) < 1e-4) cand[n++] = i;
        const int pid = cand[(s + d) % (uint64_t)n];
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
