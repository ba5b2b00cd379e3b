// Device-resident graph: CSR (out-edges) + CSC (in-edges), uint64 offsets,
// uint32 vertex ids.  At one process GRE's local graph holds exactly this
// (include/Graph/LocalGraph.hpp:23-200): edgeIndex/edgeList and
// redgeIndex/redgeList/redgeID.
#pragma once
#include "common.hpp"

#include <map>
#include <memory>
#include <utility>

namespace gre {
// Source-range partition of the CSC for XCD-local gathers.  Compact source ids
// are cut into R ranges of equal edge count; range r belongs to XCD class
// r % 8.  Entries are regrouped by (class, range, row): a (row, range) run is a
// segment, and the pull reduces segments instead of rows.  A second pull over
// each row's list of segment ids sums them in a fixed order.
// most (class, range) keys of a partition: 8 classes * (1 + 256/8)
constexpr int kMaxKeys = 264;
// Binned (propagation-blocking) layout of the forward pull; see binned.hip.
// Rows are relabelled: rows [0, nd) have in-edges and are cut into bins of
// consecutive rows; rows [nd, nV) have none.  The scatter pass walks chunks of
// consecutive source rows, the accumulate pass slices of each bin's entries.
constexpr int kBinRows = 16384;        // rows per bin (64-bit LDS accumulators)
constexpr int kBinMaxBins = 4096;      // bins (the scatter pass keeps a run table of this many in LDS)
constexpr int kEngBinRows = 8192;      // rows per bin of the engine's FP64 layout (two 64-bit LDS accumulators each)
constexpr int kChunkRows = 16384;      // source rows whose values a chunk stages in LDS
constexpr uint32_t kChunkEdges = 262144;   // entries per chunk (a longer source row is cut)
constexpr uint32_t kSliceEdges = 1u << 20;  // entries per accumulate slice (a longer bin is cut)
// the scatter's chunk stretches, one per XCD: XCD x runs chunks [s[x], s[x+1])
struct XcdStretches {
  uint32_t s[9];
};
struct BinPlan {
  // Chunks start at 64-entry boundaries of the word array and bins at 64-entry
  // boundaries of the bin array, so every wave's loads cover whole lines; the
  // gaps are never scattered into and hold zero contributions.
  uint64_t E = 0;      // entries
  uint64_t Ew = 0;     // words, with each chunk's gap
  uint64_t Eb = 0;     // bin-array slots, with each bin's gap
  uint64_t ngrp = 0;   // groups (one slot each): entries of one chunk to one destination, cut at 16-word rows
  uint32_t bin_rows = 0;  // rows per bin (kBinRows, or kEngBinRows for the engine's layout)
  uint32_t chunk_rows = 0;  // source rows per scatter chunk
  uint32_t max_in = 0;    // the engine's layout: largest in-degree
  DevBuf rowv;     // the engine's layout: uint32 [nd] vertex of each row (ascending)
  uint32_t nd = 0, B = 0, nch = 0, nsl = 0, nmb = 0, nmrows = 0;
  XcdStretches xs{};    // chunk stretches of the XCDs, cut at equal entry counts
  uint32_t xs_max = 0;  // the longest stretch (the scatter's grid is 8 of these)
  DevBuf ch_r0;    // uint32 [nch] first source row of chunk w
  DevBuf ch_nr;    // uint32 [nch] source rows chunk w stages
  DevBuf ch_e0;    // uint32 [nch+1] entries before chunk w (chunk w has ch_e0[w+1] - ch_e0[w])
  DevBuf ch_w0;    // uint32 [nch] first word of chunk w (a multiple of 64)
  // uint16 [Ew] scatter order (by bin, then destination): source row - ch_r0
  // (14 bits) | run start << 14 | group start << 15
  DevBuf src16;
  // per 64 words: (runs of the chunk started before word 64k + bit 0 of the
  // block's run starts - 1) | (groups of the chunk before word 64k) << 12
  DevBuf rcnt;     // uint32 [Ew/64]
  DevBuf ch_q0;    // uint32 [nch+1] first run of chunk w in rbase
  DevBuf rbase;    // uint32 [runs] first slot of each run minus the chunk's groups before it (mod 2^32)
  DevBuf wpos;     // uint32 [Ew] CSR position of each word's entry (edge weights follow it; 0 in the gaps)
  DevBuf dloc;     // uint16 [Eb] each slot's destination row - bin_r0[bin] (0 in the gaps)
  DevBuf bin_r0;   // uint32 [B+1] first row of each bin
  DevBuf sl_e0;    // uint32 [nsl+1] slice ranges in the bin array (a bin's last slice takes its gap)
  DevBuf sl_bin;   // uint32 [nsl] bin of each slice
  DevBuf sl_g;     // uint32 [nsl] ~0: the slice's bin has one slice; else the bin's first slot in the multi-slice sums
  DevBuf mb_r0;    // uint32 [nmb] first row of each multi-slice bin
  DevBuf mb_g;     // uint32 [nmb+1] its first slot in the multi-slice sums (mb_g[nmb] = nmrows)
};
struct XPart {
  int R = 0;
  uint32_t H = 0;    // sources [0, H) form one range shared by all XCDs (cut by rows)
  uint64_t nseg = 0;
  DevBuf idx;        // uint32 [E]   compact source ids, grouped by (class, range, row)
  DevBuf perm;       // uint32 [E]   CSC position of each regrouped entry (edge weights follow it)
  DevBuf seg_off;    // uint32 [nseg+1] (E < 2^32)
  DevBuf pair_off;   // uint32 [nV+1] segments of row v: pair_list[pair_off[v] .. pair_off[v+1])
  DevBuf pair_list;  // uint32 [nseg] segment ids, by row then class-major range
  uint64_t class_start[9] = {0};  // first entry of each class; class_start[8] = E
  // Rows relabelled (CSC partitions for the PageRank forward): partition row r
  // is vertex rowinv[r], rowperm[v] = r, so hub rows do not cluster in the
  // first tiles; outdeg_r / cid_r hold the vertices' out-degree and compact id
  // in row order.  Empty when rows are the vertices themselves.
  bool permuted = false;
  DevBuf rowperm, rowinv, outdeg_r, cid_r;
  // set for the binned layout (graph_binpart); the segment fields stay empty
  bool binned = false;
  BinPlan bin;
  int K = 0;         // (class, range) keys: 8 * (1 + R/8)
  DevBuf kseg;       // uint32 [K+1] first segment of each key; kseg[K] = nseg
  struct Tiles {
    DevBuf rb;       // uint32 [ntiles+1] tile row bounds over segments
    DevBuf xmap;     // uint32 [9] first tile of each class; xmap[8] = ntiles
    uint64_t ntiles = 0;
    unsigned grid = 0;  // 8 * (largest class tile count)
  };
  std::map<uint32_t, Tiles> tiles;  // by tile edges
  // Tiles of the per-row pass.  A tile's segments are, per key, one run of
  // consecutive segment ids (segments of a key are ordered by row), so the
  // tile stages its runs into LDS with coalesced loads and reads each pair
  // entry's sum from there.
  struct PairTiles {
    DevBuf rb;       // uint32 [ntiles+1] tile row bounds over the pair lists
    // A tile's LDS stage holds its runs in key order.  words[t*W + w] (W =
    // tile entries / 32) = (bit b set if a non-empty run starts at stage entry
    // 32w + b, number of non-empty runs starting before entry 32w); base[t*K
    // + r] = first segment of the tile's r-th non-empty run minus its stage
    // offset, so stage entry i holds segment base[rank(i)] + i.
    DevBuf words;    // uint2 [ntiles*W]
    DevBuf base;     // uint32 [ntiles*K]
    DevBuf slot;     // uint16 [ntiles*te] LDS slot of each pair entry's segment sum (zero-padded)
    uint64_t ntiles = 0;
  };
  std::map<uint32_t, PairTiles> pair_tiles;  // by tile entries
};
// Vertices whose engine pull rows (in-edges, out-edges or both) are too long
// for one thread: mid = one wave per row, lng = one block per row.
struct PullRows {
  DevBuf mid, lng;
  uint32_t n_mid = 0, n_long = 0;
};
// Rows of one adjacency (CSC or CSR) longer than a threshold, cut into chunks
// of consecutive entries (one wave each) for the per-vertex traversal kernels;
// chunks are listed row by row, in entry order.
struct LongRows {
  DevBuf rows;       // uint32 [n] vertex ids
  DevBuf cv;         // uint32 [nchunks] index into rows
  DevBuf cb;         // uint64 [nchunks] first entry of the chunk
  DevBuf cstart;     // uint32 [n+1] first chunk of each row
  DevBuf acc;        // uint64 [n] per-row accumulator of the users (zero between uses)
  uint32_t n = 0, nchunks = 0;
};
}  // namespace gre

struct gre_graph {
  uint32_t nV = 0;
  uint64_t nE = 0;
  gre::DevBuf off, idx;     // CSR
  gre::DevBuf coff, cidx;   // CSC
  gre::DevBuf ceid;         // CSC position -> CSR position (uint64), built on demand
  gre::DevBuf outdeg;       // int32
  gre::DevBuf w_u32;        // SSSP weights, CSR order (optional)
  gre::DevBuf w_u32_csc;    // same, CSC order (built on demand)
  bool rows_sorted = false; // CSR rows sorted by destination (R-MAT, from_csr sorted)
  // tile row partitions of the pull kernels, keyed by (csc?, tile edges)
  std::map<std::pair<int, uint32_t>, gre::DevBuf> tiles;
  // hub-sorted compact source ids (built on demand by the pull schedules):
  // cid[v] = rank of v among vertices with out-edges, by descending out-degree;
  // cidx_c[p] = cid[cidx[p]] (CSC in compact source ids)
  gre::DevBuf cid, cidx_c;
  uint32_t nsrc = 0;
  // did[v] = rank of v among vertices with in-edges, by descending in-degree;
  // idx_c[p] = did[idx[p]] (CSR in compact target ids)
  gre::DevBuf did, idx_c;
  uint32_t ntgt = 0;
  // source-range partitions, by (tail ranges, shared sources) (built on demand, shared with
  // handles); over the CSC by (R, H), with relabelled rows by (R | 1024, H), over the CSR (compact
  // targets) by (-R, H)
  std::map<std::pair<int, uint32_t>, std::shared_ptr<gre::XPart>> xpart;
  // engine pull row classes by direction (built on demand)
  std::map<int, gre::PullRows> pull_rows;
  std::map<uint64_t, gre::LongRows> long_rows;  // by (csc, threshold, chunk entries)
  // long-row (threshold, chunk) pairs of the 64-source BFS's bottom-up pass and
  // of the BC forward pass: 0 until the first call times the candidates (or the
  // caller sets them)
  uint64_t ms_long = 0, ms_chunk = 0;
  uint64_t bc_flong = 0, bc_fchunk = 0;
  // the reversed graph (built on demand): its CSR and CSC are views of this
  // graph's CSC and CSR, its out-degrees are this graph's in-degrees
  std::unique_ptr<gre_graph> rev;
};

namespace gre {
// CSC -> CSR position map (built once, cached on the graph).
const uint64_t* graph_ceid(gre_graph* g);
// SSSP weights permuted to CSC order (cached).
const uint32_t* graph_w_csc(gre_graph* g);
// tile_rb[t] = #rows r < nV with off[r] < t*te, t in [0, ntiles]; ntiles >= 1.
const uint32_t* graph_tiles(gre_graph* g, bool csc, uint32_t te, uint64_t* ntiles);
// Builds g->cid / g->cidx_c / g->nsrc once (hub-sorted compact source ids).
void graph_compact_sources(gre_graph* g);
// Builds g->did / g->idx_c / g->ntgt once (hub-sorted compact target ids).
void graph_compact_targets(gre_graph* g);
// tile_rb over an arbitrary offset array of nrows rows (see graph_tiles).
void build_tile_rows(const uint64_t* off, uint32_t nrows, uint64_t ntiles, uint32_t te, uint32_t* rb,
                     hipStream_t s);
void build_tile_rows(const uint32_t* off, uint32_t nrows, uint64_t ntiles, uint32_t te, uint32_t* rb,
                     hipStream_t s);
// The source partition of g: the H hottest compact sources as one shared range
// and the rest in R tail ranges (R in {8, 16, ..., 256}); built once and cached.
// csc = false partitions the CSR the same way, over compact target ids.
// rows_permuted (csc only): the partition's rows are the vertices relabelled by
// a fixed bijection (XPart::permuted); cached apart from the plain one.
std::shared_ptr<XPart> graph_xpart(gre_graph* g, int R, uint32_t H, bool csc = true, bool rows_permuted = false);
// The binned layout of g's forward pull (E < 2^32, at most kBinMaxBins bins);
// built once and cached with the partitions.
std::shared_ptr<XPart> graph_binpart(gre_graph* g);
// The binned layout of the engine's FP64 sum pull (binned.hip), cached on the
// graph like the other partitions.
std::shared_ptr<XPart> graph_binpart_engine(gre_graph* g);
// The reversed graph of g (cached on g; the views stay valid while g lives).
gre_graph* graph_reversed(gre_graph* g);
// Drops the cached partitions that no PageRank handle or engine holds any more
// (the reversed graph's too).
void prune_xparts(gre_graph* g);
// Its tiles of te entries (cached on the partition).
const XPart::Tiles& xpart_tiles(XPart& xp, uint32_t te, hipStream_t s);
// Tiles of te entries over the per-row pair lists (rows = vertices).
const XPart::PairTiles& xpart_pair_tiles(XPart& xp, uint32_t nV, uint32_t te, hipStream_t s);

// device helpers
__device__ inline uint32_t row_of(const uint64_t* __restrict__ off, uint32_t nV, uint64_t e) {
  // largest r with off[r] <= e (rows with off[r] == off[r+1] are skipped)
  uint32_t lo = 0, hi = nV;  // answer in [lo, hi)
  while (hi - lo > 1) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (off[mid] <= e) lo = mid; else hi = mid;
  }
  return lo;
}

// exclusive scan helpers (rocPRIM), sizes in elements
void scan_excl_u64(const uint64_t* in, uint64_t* out, size_t n, hipStream_t s);
void scan_excl_u32_to_u64(const uint32_t* in, uint64_t* out, size_t n, hipStream_t s);
}  // namespace gre
