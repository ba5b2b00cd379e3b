// The vertex-program base class users derive from
// (GRE::VertexCode<S,C,V,DG>, include/VertexCode/VertexCode.hpp:9-212 in the
// GRE tree).  Same interface: the primitives a program overrides (init,
// scatter, combine, apply, assert_to_halt, reduce, output_statistics) and the
// array helpers (allocScatterData & co) that size and record the program's
// scatter/combine/vertex arrays.  SynchronousEngine reads the recorded
// pointers to move the arrays between host memory and HBM around each run.
#ifndef GRE_AMD_VERTEXCODE_HPP
#define GRE_AMD_VERTEXCODE_HPP
#include <cassert>
#include <iostream>

#include "Graph/DistributedGraph.hpp"
#include "Message.hpp"
#include "Types.hpp"
#include "Util/Util.hpp"

namespace GRE {
template <typename VC> class SynchronousEngine;

template <typename scatter_data_t_, typename combine_data_t_, typename vertex_data_t_, typename DistributedGraph_t>
class VertexCode {
 public:
  typedef scatter_data_t_ scatter_data_t;
  typedef combine_data_t_ combine_data_t;
  typedef vertex_data_t_ vertex_data_t;
  typedef scatter_data_t ScatterData_t;
  typedef combine_data_t CombineData_t;
  typedef vertex_data_t VertexData_t;
  template <typename VC> friend class GRE::SynchronousEngine;

  // ---- primitives (defaults report the missing override, like GRE's) ----
  bool init(const lvid_t) {
    std::cerr << "Fetal Error! Init primitive is not implemented." << std::endl;
    return false;
  }
  template <typename Ctx_t, typename EdgeIterator_t>
  void scatter(Ctx_t&, const lvid_t, const lvid_t, const EdgeIterator_t&) {
    std::cerr << "Fetal Error! Scatter primitive is not implemented." << std::endl;
  }
  template <typename Lock_t>
  bool combine(const lvid_t, const combine_data_t&, Lock_t&) {
    std::cerr << "Fetal Error! Combine primitive is not implemented." << std::endl;
    return false;
  }
  bool apply(const lvid_t) {
    std::cerr << "Fetal Error! Apply primitive is not implemented." << std::endl;
    return false;
  }
  bool assert_to_halt(const lvid_t) { return true; }
  void aggregate(lvid_t) {}
  template <typename DC_t>
  void reduce(DC_t&) {}
  template <typename DC_t>
  void output_statistics(DC_t&) {}

  // ---- data movement helpers (GRE's defaults) ----
  template <typename Ctx_t>
  void copy_scatter_data(Ctx_t&, scatter_data_t& d, const lvid_t s) { d = scatter_data[s]; }
  void copy_scatter_data(scatter_data_t& d, const lvid_t s) { d = scatter_data[s]; }
  template <typename Ctx_t>
  void update_scatter(Ctx_t&, const lvid_t v, const scatter_data_t& d) { scatter_data[v] = d; }
  void update_scatter(const lvid_t v, const scatter_data_t& d) { scatter_data[v] = d; }
  template <typename Ctx_t>
  void copy_combine_data(Ctx_t&, combine_data_t& d, const lvid_t v) { d = combine_data[v]; }
  void copy_combine_data(combine_data_t& d, const lvid_t v) { d = combine_data[v]; }
  template <typename Ctx_t>
  void reset_combiner(Ctx_t&, const lvid_t) {
    std::cerr << "Fetal Error! reset combiner primitive is not implemented." << std::endl;
  }
  void reset_combiner(const lvid_t) {
    std::cerr << "Fetal Error! reset combiner primitive  is not implemented." << std::endl;
  }
  template <typename Ctx_t>
  void reset_scatter(Ctx_t&, const lvid_t) {}
  void reset_scatter(const lvid_t) {}

 protected:
  scatter_data_t* scatter_data;
  combine_data_t* combine_data;
  vertex_data_t* vertex_data;
  DistributedGraph_t& graph;

  explicit VertexCode(DistributedGraph_t& g) : scatter_data(NULL), combine_data(NULL), vertex_data(NULL), graph(g) {}

 public:
  ~VertexCode() {}

  // scatter/combine arrays: align(masters) + agents entries; vertex arrays: masters
  void allocScatterData(scatter_data_t*& p) {
    assert(!p);
    p = new scatter_data_t[scatterCount()];
    scatter_data = p;
  }
  void allocScatterData(scatter_data_t*& p, const scatter_data_t iv) {
    allocScatterData(p);
    fill(p, scatterCount(), iv);
  }
  void allocCombineData(combine_data_t*& p) {
    assert(!p);
    p = new combine_data_t[combineCount()];
    combine_data = p;
  }
  void allocCombineData(combine_data_t*& p, const combine_data_t iv) {
    allocCombineData(p);
    fill(p, combineCount(), iv);
  }
  void allocVertexData(vertex_data_t*& p) {
    if (!p) p = new vertex_data_t[graph.get_num_masters()];
    vertex_data = p;
  }
  void allocVertexData(vertex_data_t*& p, const vertex_data_t iv) {
    allocVertexData(p);
    fill(p, (long)graph.get_num_masters(), iv);
  }
  void setScatterData(const scatter_data_t* p) { scatter_data = const_cast<scatter_data_t*>(p); }
  void setCombineData(const combine_data_t* p) { combine_data = const_cast<combine_data_t*>(p); }
  void setVertexData(const vertex_data_t* p) { vertex_data = const_cast<vertex_data_t*>(p); }
  void freeScatterData(scatter_data_t*& p) {
    delete[] p;
    p = NULL;
    scatter_data = NULL;
  }
  void freeCombineData(combine_data_t*& p) {
    delete[] p;
    p = NULL;
    combine_data = NULL;
  }
  void freeVertexData(vertex_data_t*& p) {
    delete[] p;
    p = NULL;
    vertex_data = NULL;
  }

 private:
  long scatterCount() { return (long)(Util::align(graph.get_num_masters()) + graph.get_num_scatters()); }
  long combineCount() { return (long)(Util::align(graph.get_num_masters()) + graph.get_num_combiners()); }
  template <typename T>
  static void fill(T* p, long n, const T& v) {
    for (long i = 0; i < n; ++i) p[i] = v;
  }
};

}  // namespace GRE
#endif
