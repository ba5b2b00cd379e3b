// Host view of the local graph (GraphInCSR, include/Graph/LocalGraph.hpp:23-200
// in the GRE tree).  The graph itself lives in HBM (gre_graph, built once at
// ingress); this class downloads CSR and CSC on first host access, for code
// that walks edges on the host.  The device superstep never uses it.
//
// Deviation from GRE, kept on purpose: leid(it) is the CSR position of the
// edge (also for an iterator of the reverse list).  GRE divides that position by sizeof(lidx_t) (LocalGraph.hpp:173-176),
// so edge property e was read for edge 8e; and GRE's redge_begin/end index the
// CSR offsets (:183-191), which is right only when every in-degree equals the
// out-degree.  Here redge_* use the CSC offsets and rleid maps a CSC position
// to the CSR position of the same edge.
#ifndef GRE_AMD_LOCALGRAPH_HPP
#define GRE_AMD_LOCALGRAPH_HPP
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "Types.hpp"
#include "gre_dev.h"

namespace GRE {
namespace Graph {

template <typename T> class DistributedGraph;
template <typename T> class GraphIngress;

class GraphInCSR {
 public:
  typedef std::vector<lidx_t> EdgeIndex_t;
  typedef std::vector<lvid_t> EdgeList_t;
  typedef EdgeList_t::iterator edge_iterator;
  typedef EdgeList_t::iterator redge_iterator;

  GraphInCSR() : dev_(NULL), nVertices_(0), nEdges_(0), host_(false) {}

  edge_iterator edge_begin() { ensure(); return edgeList_.begin(); }
  edge_iterator edge_end() { ensure(); return edgeList_.end(); }
  edge_iterator edge_begin(lvid_t v) { ensure(); return edgeList_.begin() + edgeIndex_[v]; }
  edge_iterator edge_end(lvid_t v) { ensure(); return edgeList_.begin() + edgeIndex_[v + 1]; }
  // the CSR position of the edge; an iterator into the reverse list (a
  // program scattering along In) is mapped through the CSC -> CSR table
  leid_t leid(const edge_iterator& it) {
    ensure();
    const lvid_t* p = &*it;
    if (nEdges_ > 0 && p >= redgeList_.data() && p < redgeList_.data() + nEdges_)
      return (leid_t)redgeID_[p - redgeList_.data()];
    return (leid_t)(it - edgeList_.begin());
  }

  redge_iterator redge_begin() { ensure(); return redgeList_.begin(); }
  redge_iterator redge_end() { ensure(); return redgeList_.end(); }
  redge_iterator redge_begin(lvid_t v) { ensure(); return redgeList_.begin() + redgeIndex_[v]; }
  redge_iterator redge_end(lvid_t v) { ensure(); return redgeList_.begin() + redgeIndex_[v + 1]; }
  leid_t rleid(const redge_iterator& it) { ensure(); return (leid_t)redgeID_[it - redgeList_.begin()]; }

  size_t get_num_edges() const { return nEdges_; }
  size_t get_num_vertices() const { return nVertices_; }

  // the device graph this view mirrors
  gre_graph* device() const { return dev_; }
  void attach(gre_graph* g, size_t nV, size_t nE) {
    dev_ = g;
    nVertices_ = nV;
    nEdges_ = nE;
    host_ = false;
  }

 private:
  void ensure() {
    if (host_) return;
    if (!dev_) throw std::runtime_error("GraphInCSR: no graph loaded (call ingress first)");
    std::vector<uint64_t> off(nVertices_ + 1), coff(nVertices_ + 1), ceid(nEdges_);
    edgeList_.resize(nEdges_);
    redgeList_.resize(nEdges_);
    const int rc = gre_graph_download(dev_, off.data(), edgeList_.data(), coff.data(), redgeList_.data(),
                                      ceid.data(), NULL);
    if (rc != 0) throw std::runtime_error(std::string("gre_graph_download: ") + gre_last_error());
    edgeIndex_.assign(off.begin(), off.end());
    redgeIndex_.assign(coff.begin(), coff.end());
    redgeID_.assign(ceid.begin(), ceid.end());
    host_ = true;
  }

  gre_graph* dev_;
  size_t nVertices_, nEdges_;
  bool host_;
  EdgeIndex_t edgeIndex_, redgeIndex_, redgeID_;
  EdgeList_t edgeList_, redgeList_;
};

}  // namespace Graph
}  // namespace GRE
#endif
