// Vertex-cut placement of GRE's multi-process ingress, for np parts:
// the LDG edge assignment each reading process makes (GRE's
// GraphIngress::LDGEdgeDecision, include/Graph/GraphIngress.hpp:1126-1171,
// driven by loadEdgesFromFiles :511-723) and the per-part agent counts and
// local id layout the ingress then builds (:111-247).  One GPU runs one part
// (all vertices masters); this is the placement a run over np GPUs would
// load, computed on the host like GRE's (the decisions are sequential per
// reader).
//
// Deviations from GRE, on purpose:
//   - GRE's per-vertex part bitmaps start uninitialised (FixedBitmap's
//     constructor does not clear them); here they start empty;
//   - fewer than np*np edges make GRE's capacity 0 and divide by it; here
//     that is an error.
#include "common.hpp"

#include <cmath>
#include <vector>

namespace gre {
namespace {

// GRE::Util::align: the master count padded to a whole 64-entry word
uint64_t align64(uint64_t n) { return (n + 63) / 64 * 64; }

uint64_t masters_of(uint32_t nV, int np, int p) {
  return nV / (uint32_t)np + ((uint32_t)p < nV % (uint32_t)np ? 1 : 0);
}

// One reader's LDG state: the parts each vertex already has as a source and as
// a destination (np <= 64 bits), and the edges given to each part.
struct Ldg {
  int np;
  double capacity;
  std::vector<uint64_t> srcParts, dstParts;
  std::vector<uint64_t> counts;
  std::vector<double> score;
  std::vector<int> stack;
  Ldg(uint32_t nV, int np_, uint64_t cap)
      : np(np_), capacity((double)cap), srcParts(nV, 0), dstParts(nV, 0), counts(np_, 0), score(np_), stack(np_) {}

  int place(uint32_t src, uint32_t dst) {
    for (int i = 0; i < np; ++i) {
      const double bal = std::fabs(1.0 - (double)counts[i] / capacity);
      const bool sOn = (srcParts[src] >> i) & 1u, dOn = (dstParts[dst] >> i) & 1u;
      const double sHome = (int)(src % (uint32_t)np) == i ? 0.5 : 0.0;
      const double dTerm = dOn ? 1.0 : ((int)(dst % (uint32_t)np) == i ? 0.5 : 0.0);
      // GRE's expression `bal * (sOn ? 1.0 : sHome + dTerm)`: a part that
      // already has the source scores bal whatever the destination
      score[i] = bal * (sOn ? 1.0 : sHome + dTerm);
    }
    double best = score[0];
    for (int i = 1; i < np; ++i)
      if (score[i] > best) best = score[i];
    int top = 0;
    for (int i = 0; i < np; ++i)
      if (std::fabs(best - score[i]) < 1e-4) stack[top++] = i;
    // GRE's tie break on (src + dst) as 64-bit vertex ids
    const int pid = stack[(int)(((uint64_t)src + (uint64_t)dst) % (uint64_t)top)];
    counts[pid]++;
    srcParts[src] |= 1ull << pid;
    dstParts[dst] |= 1ull << pid;
    return pid;
  }
};

void check_common(uint32_t nV, int np, uint64_t nE, const uint32_t* src, const uint32_t* dst) {
  GRE_CHECK(np >= 1 && np <= 64, GRE_DEV_NULL, "np must be in [1, 64] (GRE's MAX_NUM_MACHINES)");
  GRE_CHECK(nE == 0 || (src && dst), GRE_DEV_NULL, "null edge arrays");
  for (uint64_t e = 0; e < nE; ++e)
    GRE_CHECK(src[e] < nV && dst[e] < nV, GRE_DEV_NULL, "vertex id >= nV");
}

}  // namespace
}  // namespace gre

using namespace gre;

extern "C" {

int gre_vertex_cut_ldg(uint32_t nV, uint64_t nE_meta, int np, int nfiles, const uint64_t* file_off,
                       const uint32_t* src, const uint32_t* dst, int32_t* part) {
  return guard([&] {
    GRE_CHECK(nfiles >= 0 && file_off && part, GRE_DEV_NULL, "null argument");
    const uint64_t nE = file_off[nfiles];
    GRE_CHECK(file_off[0] == 0, GRE_DEV_NULL, "file_off[0] must be 0");
    for (int f = 0; f < nfiles; ++f) GRE_CHECK(file_off[f] <= file_off[f + 1], GRE_DEV_NULL, "file_off not sorted");
    check_common(nV, np, nE, src, dst);
    const uint64_t cap = nE_meta / (uint64_t)np / (uint64_t)np;   // GraphIngress.hpp:560
    GRE_CHECK(cap > 0, GRE_DEV_NULL, "fewer than np*np edges: GRE's LDG capacity would be 0");
    // reader r loads files r, r + np, ... in order, with its own decision state
    for (int r = 0; r < np; ++r) {
      Ldg ldg(nV, np, cap);
      for (int f = r; f < nfiles; f += np)
        for (uint64_t e = file_off[f]; e < file_off[f + 1]; ++e) part[e] = ldg.place(src[e], dst[e]);
    }
  });
}

int gre_vertex_cut_stats(uint32_t nV, int np, int dir, uint64_t nE, const uint32_t* src, const uint32_t* dst,
                         const int32_t* part, int64_t* stats) {
  return guard([&] {
    GRE_CHECK(stats && (nE == 0 || part), GRE_DEV_NULL, "null argument");
    GRE_CHECK(dir >= GRE_DIR_OUT && dir <= GRE_DIR_INOUT, GRE_DEV_NULL, "bad direction");
    check_common(nV, np, nE, src, dst);
    for (uint64_t e = 0; e < nE; ++e) GRE_CHECK(part[e] >= 0 && part[e] < np, GRE_DEV_NULL, "part id out of range");
    // per part: scatter agents (remote sources of its edges) and combiner
    // agents (remote destinations), GraphIngress.hpp:111-149
    for (int p = 0; p < np; ++p) {
      std::vector<uint8_t> sc(nV, 0), cb(nV, 0);
      int64_t nS = 0, nC = 0, nB = 0, nEp = 0;
      for (uint64_t e = 0; e < nE; ++e) {
        if (part[e] != p) continue;
        ++nEp;
        if ((int)(src[e] % (uint32_t)np) != p && !sc[src[e]]) { sc[src[e]] = 1; ++nS; }
        if ((int)(dst[e] % (uint32_t)np) != p && !cb[dst[e]]) { cb[dst[e]] = 1; ++nC; }
      }
      for (uint32_t v = 0; v < nV; ++v) nB += sc[v] && cb[v];
      const int64_t agents = nC + nS - nB;
      int64_t* o = stats + 7 * (int64_t)p;
      o[0] = (int64_t)masters_of(nV, np, p);
      // In / InOut: both agent kinds become one agent set (:150-153)
      o[1] = dir == GRE_DIR_OUT ? nS : agents;
      o[2] = dir == GRE_DIR_OUT ? nC : agents;
      o[3] = nB;
      o[4] = agents;
      o[5] = nEp;
      o[6] = (int64_t)align64((uint64_t)o[0]);
    }
  });
}

int gre_vertex_cut_lvid2vid(uint32_t nV, int np, int dir, int p, uint64_t nE, const uint32_t* src,
                            const uint32_t* dst, const int32_t* part, uint64_t* lvid2vid, uint64_t cap,
                            uint64_t* n_out) {
  return guard([&] {
    GRE_CHECK(n_out && (nE == 0 || part), GRE_DEV_NULL, "null argument");
    GRE_CHECK(dir >= GRE_DIR_OUT && dir <= GRE_DIR_INOUT, GRE_DEV_NULL, "bad direction");
    check_common(nV, np, nE, src, dst);
    GRE_CHECK(p >= 0 && p < np, GRE_DEV_NULL, "part out of range");
    std::vector<uint8_t> sc(nV, 0), cb(nV, 0);
    for (uint64_t e = 0; e < nE; ++e) {
      if (part[e] != p) continue;
      if ((int)(src[e] % (uint32_t)np) != p) sc[src[e]] = 1;
      if ((int)(dst[e] % (uint32_t)np) != p) cb[dst[e]] = 1;
    }
    // masters p, p + np, ...; padding to a whole word with -1; then the
    // combiners (Out) or all agents (In, InOut) in ascending id (:154-178)
    std::vector<uint64_t> out;
    for (uint64_t v = (uint64_t)p; v < nV; v += (uint64_t)np) out.push_back(v);
    const uint64_t nm = out.size();
    for (uint64_t i = nm; i < align64(nm); ++i) out.push_back(~0ull);
    for (uint32_t v = 0; v < nV; ++v)
      if (cb[v] || (dir != GRE_DIR_OUT && sc[v])) out.push_back(v);
    *n_out = out.size();
    if (lvid2vid) {
      GRE_CHECK(cap >= out.size(), GRE_DEV_NOBUFFER, "lvid2vid buffer too small");
      for (size_t i = 0; i < out.size(); ++i) lvid2vid[i] = out[i];
    }
  });
}

}  // extern "C"
