// Graph ingress (include/Graph/GraphIngress.hpp:24-304,511-723 in the GRE tree)
// at one process: <prefix>.meta holds "nVertices nEdges"; <prefix>.edges.0,
// .edges.1, ... hold one "src dst" pair per line ('#' lines and blank lines are
// skipped, ids parsed with strtoul as GRE does).  Edges keep file order, which
// fixes the in-row order of the CSR and CSC built on the device
// (gre_graph_from_edges).  Format "bin" reads the same files with stream
// extraction (GRE's loadEdgesFromBinaryFiles reads text the same way).
// The "partition" method (offline pre-partitioned files) is not supported.
#ifndef GRE_AMD_GRAPH_INGRESS_HPP
#define GRE_AMD_GRAPH_INGRESS_HPP
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "Types.hpp"
#include "Util/Util.hpp"

namespace GRE {
namespace Graph {

struct IngressParameters {
  std::string format;
  std::string method;
  std::string prefix;
  IngressParameters() : format("ascii"), method("edgelists"), prefix() {}
  explicit IngressParameters(std::string p) : format("ascii"), method("edgelists"), prefix(p) {}
  IngressParameters(std::string p, std::string f, std::string m) : format(f), method(m), prefix(p) {}
  void setFormat(std::string f) { format = (f == "bin") ? f : "ascii"; }
  void setMethod(std::string& m) { method = (m == "partitions") ? m : "edgelists"; }
  void setPrefix(std::string& p) { prefix = p; }
  std::string& getPrefix() { return prefix; }
  bool isBinary() const { return format == "bin"; }
  bool isAscii() const { return format != "bin"; }
  bool isEdgelists() const { return method != "partition" && method != "partitions"; }
  bool isPartitions() const { return !isEdgelists(); }
};

// Reads the edge files into (src, dst) arrays in file order and counts
// out-degrees while reading, as GRE does (GraphIngress.hpp:581-591).
struct EdgeFileReader {
  static bool readMeta(const std::string& prefix, size_t& nV, size_t& nE) {
    std::ifstream in((prefix + ".meta").c_str());
    if (!in) {
      std::cerr << "Err: cannot open " << prefix << ".meta" << std::endl;
      return false;
    }
    in >> nV >> nE;
    return !in.fail();
  }
  static bool readEdges(const std::string& prefix, bool binary, size_t nV, std::vector<uint32_t>& src,
                        std::vector<uint32_t>& dst, std::vector<int>& outdeg) {
    outdeg.assign(nV, 0);
    for (int i = 0;; ++i) {
      const std::string fn = prefix + ".edges." + Util::toString(i);
      std::ifstream in(fn.c_str());
      if (!in) break;
      if (binary) {
        vid_t s, d;
        while (in >> s >> d) {
          if (!push(s, d, nV, src, dst, outdeg, fn)) return false;
        }
        continue;
      }
      std::string line;
      while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        char* rest = NULL;
        const vid_t s = std::strtoul(line.c_str(), &rest, 10);
        const vid_t d = std::strtoul(rest, NULL, 10);
        if (!push(s, d, nV, src, dst, outdeg, fn)) return false;
      }
    }
    return true;
  }

 private:
  static bool push(vid_t s, vid_t d, size_t nV, std::vector<uint32_t>& src, std::vector<uint32_t>& dst,
                   std::vector<int>& outdeg, const std::string& fn) {
    if (s >= nV || d >= nV) {
      std::cerr << "Err: edge (" << s << "," << d << ") in " << fn << " has a vertex id >= " << nV << std::endl;
      return false;
    }
    src.push_back((uint32_t)s);
    dst.push_back((uint32_t)d);
    outdeg[s]++;
    return true;
  }
};

// GRE's ingress object; DistributedGraph::ingress drives it.
template <typename DistributedGraph_t>
class GraphIngress {
 public:
  explicit GraphIngress(DistributedGraph_t& g) : g_(g) {}
  template <Dir dir>
  bool ingress(IngressParameters& p) {
    if (p.isPartitions()) {
      std::cerr << "Error: the \"partition\" ingress method is not supported on one GPU." << std::endl;
      return false;
    }
    size_t nV = 0, nE = 0;
    if (!EdgeFileReader::readMeta(p.prefix, nV, nE)) return false;
    std::vector<uint32_t> src, dst;
    std::vector<int> outdeg;
    if (!EdgeFileReader::readEdges(p.prefix, p.isBinary(), nV, src, dst, outdeg)) return false;
    return g_.adoptEdges(nV, src, dst, outdeg);
  }
  bool finalize() { return true; }

 private:
  DistributedGraph_t& g_;
};

}  // namespace Graph
}  // namespace GRE
#endif
