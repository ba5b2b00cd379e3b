// Frontier sets (include/Graph/VertexSet.hpp:17-347 in the GRE tree).
//
// SourceVertexSet is the run() argument: setALL() or a list of source lvids.
// The engine's own frontiers (GRE's currSet/nextSet: a bitmap plus a vertex
// vector) live in HBM inside the device engine; VertexSet here is a host
// implementation of the same interface for host-side callers.
#ifndef GRE_AMD_VERTEXSET_HPP
#define GRE_AMD_VERTEXSET_HPP
#include <stdint.h>

#include <fstream>
#include <string>
#include <vector>

#include "Types.hpp"
#include "Util/Bitmap.hpp"
#include "Util/Util.hpp"

namespace GRE {
namespace Graph {

class VertexSet {
 public:
  VertexSet() : begin_(0), end_(0), end1_(0), capacity_(0), initialized_(false) {}
  explicit VertexSet(int capacity, bool use_set = true)
      : begin_(0), end_(0), end1_(0), capacity_(0), initialized_(false) {
    initialize(capacity, use_set);
  }
  template <typename DistributedGraph_t>
  VertexSet(DistributedGraph_t& g, bool use_set = true)
      : begin_(0), end_(0), end1_(0), capacity_(0), initialized_(false) {
    initialize(g, use_set);
  }
  void initialize(int capacity, bool use_set = true) {
    if (initialized_) return;
    capacity_ = capacity;
    bits_.resize(capacity_);
    if (use_set) set_.assign(capacity_, (lvid_t)-1);
    begin_ = end_ = end1_ = 0;
    initialized_ = true;
  }
  template <typename DistributedGraph_t>
  void initialize(DistributedGraph_t& g, bool use_set = true) {
    initialize((int)(Util::align(g.get_num_masters()) + g.get_num_scatters() + g.get_num_combiners()), use_set);
  }
  void destroy() {
    bits_.destroy();
    set_.clear();
    initialized_ = false;
  }

  void set(lvid_t v) { bits_.set(v); }
  void unset(lvid_t v) { bits_.unset(v); }
  bool test(lvid_t v) { return bits_.test(v); }
  bool testAndSet(lvid_t v) { return bits_.testAndSet(v); }

  void insert(const lvid_t* buf, int count) {
    const int64_t b = __sync_fetch_and_add(&end_, (int64_t)count);
    for (int i = 0; i < count; ++i) set_[b + i] = buf[i];
  }
  void insert(std::vector<lvid_t>& buf) { insert(buf.empty() ? NULL : &buf[0], (int)buf.size()); }
  void insert(lvid_t v) { set_[__sync_fetch_and_add(&end_, (int64_t)1)] = v; }
  void setAndInsert(lvid_t v) {
    if (!bits_.testAndSet(v)) insert(v);
  }
  void setAndInsert(lvid_t* buf, int count) {
    for (int i = 0; i < count; ++i) setAndInsert(buf[i]);
  }
  void setAndInsert(std::vector<lvid_t>& buf) {
    for (size_t i = 0; i < buf.size(); ++i) setAndInsert(buf[i]);
  }
  // dynamic work queue: hands out up to count vertices published by refresh()
  int fetch(lvid_t* buf, int count) {
    const int64_t b = __sync_fetch_and_add(&begin_, (int64_t)count);
    const int64_t left = end1_ - b;
    if (left <= 0) return 0;
    const int n = (int)(left < count ? left : count);
    for (int i = 0; i < n; ++i) buf[i] = set_[b + i];
    return n;
  }
  void refresh() {
    begin_ = 0;
    end1_ = end_;
    __sync_synchronize();
  }
  void clear() {
    begin_ = end_ = end1_ = 0;
    bits_.clear();
    __sync_synchronize();
  }
  void clear_set() {
    begin_ = end_ = end1_ = 0;
    __sync_synchronize();
  }
  void clear_bitmap() { bits_.clear(); }
  void reset() { clear(); }
  int64_t size() { return end_; }
  int64_t num_ready() { return end1_; }
  int64_t num_new() { return end_ - end1_; }
  int64_t num_left() { return end1_ - begin_; }
  bool empty() { return begin_ >= end1_; }
  int64_t capacity() { return capacity_; }
  Util::Bitmap& exportBitset() { return bits_; }
  std::vector<lvid_t>& exportSet() { return set_; }

 private:
  int64_t begin_, end_, end1_, capacity_;
  bool initialized_;
  Util::Bitmap bits_;
  std::vector<lvid_t> set_;
};

struct SourceVertexSet {
  bool isComplete;
  std::vector<lvid_t> vertices;
  SourceVertexSet() : isComplete(false) {}
  bool isNULL() { return !isComplete && vertices.empty(); }
  void setNULL() { clear(); }
  bool isALL() { return isComplete; }
  void setALL() { isComplete = true; }
  void clear() {
    isComplete = false;
    vertices.clear();
  }
  void reset() { clear(); }
  // "<prefix>.roots": a count, then that many vertex ids (include/Graph/Doc_SourceSet)
  template <typename DistributedGraph>
  void loadFromFile(const std::string& prefix, DistributedGraph& g) {
    std::ifstream in((prefix + ".roots").c_str());
    int64_t n = 0;
    in >> n;
    for (int64_t i = 0; i < n; ++i) {
      vid_t v;
      in >> v;
      if (g.is_local(v)) vertices.push_back(g.get_lvid_master(v));
    }
  }
  template <typename DistributedGraph>
  void addVertex(vid_t v, DistributedGraph& g) {
    if (g.is_local(v)) vertices.push_back(g.get_lvid_master(v));
  }
  template <typename DistributedGraph>
  void addVertices(std::vector<vid_t>& src, DistributedGraph& g, const size_t n = 0) {
    const size_t c = (n == 0 || n > src.size()) ? src.size() : n;
    for (size_t i = 0; i < c; ++i) addVertex(src[i], g);
  }
};

}  // namespace Graph
}  // namespace GRE
#endif
