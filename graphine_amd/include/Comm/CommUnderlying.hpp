// GRE's process-level communicator (COMM::Underlying,
// include/Comm/CommUnderlying.hpp:11-253 in the GRE tree) at one process.
//
// The superstep runs on one GPU, so there is no peer: reduce/allReduce copy
// the local value, barrier is a no-op, and the wall timer is std::chrono.
// Vertex programs pass MPI datatype/op constants to reduce(); when no MPI
// header is in the translation unit those names are defined here as plain
// enumerators so the programs compile unchanged.
#ifndef GRE_AMD_COMM_UNDERLYING_HPP
#define GRE_AMD_COMM_UNDERLYING_HPP
#include <chrono>
#include <cstring>
#include <iostream>
#include <vector>

#include "Types.hpp"

#ifndef MPI_VERSION
typedef int MPI_Datatype;
typedef int MPI_Op;
typedef int MPI_Comm;
enum GRE_OneProcessMPIType {
  MPI_CHAR = 1, MPI_BYTE, MPI_SHORT, MPI_INT, MPI_UNSIGNED, MPI_LONG, MPI_UNSIGNED_LONG,
  MPI_LONG_LONG, MPI_UNSIGNED_LONG_LONG, MPI_FLOAT, MPI_DOUBLE
};
enum GRE_OneProcessMPIOp { MPI_SUM = 1, MPI_MAX, MPI_MIN, MPI_PROD };
#define MPI_COMM_WORLD 0
#endif

namespace GRE {
namespace COMM {

class Underlying {
 public:
  typedef MPI_Op Op_t;
  typedef MPI_Datatype DataType_t;

  explicit Underlying(MPI_Comm c = MPI_COMM_WORLD) : comm_(c), initialized_(false), t0_(0.0) {}
  void init(int& argc, char**& argv) {
    (void)argc;
    (void)argv;
    initialized_ = true;
  }
  void finalize() { initialized_ = false; }
  int size() const { return 1; }
  int rank() const { return 0; }
  void barrier() {}
  void allReduce(void* send, void* recv, int count, DataType_t t, Op_t op) {
    (void)op;
    (void)count;
    std::memcpy(recv, send, bytes(t));  // GRE's allReduce always moves one element
  }
  int reduce(void* send, void* recv, int count, DataType_t t, Op_t op, int root) {
    (void)op;
    (void)root;
    std::memcpy(recv, send, bytes(t) * (size_t)count);
    return 0;
  }
  int voteToHalt(bool active = false) { return active ? 1 : 0; }
  void startTimer() { t0_ = now(); }
  double stopTimer() {
    const double t = now();
    const double d = t - t0_;
    t0_ = t;
    return d;
  }
  void initProfiling() {}
  void outputProfilingData(int root = 0) { (void)root; }

 private:
  static double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  }
  static size_t bytes(DataType_t t) {
    switch (t) {
      case MPI_CHAR: case MPI_BYTE: return 1;
      case MPI_SHORT: return 2;
      case MPI_INT: case MPI_UNSIGNED: case MPI_FLOAT: return 4;
      default: return 8;
    }
  }
  MPI_Comm comm_;
  bool initialized_;
  double t0_;
};

}  // namespace COMM
}  // namespace GRE
#endif
