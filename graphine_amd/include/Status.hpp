// GRE_Status and its parser (include/Status.hpp:5-34 in the GRE tree); the
// C ABI returns these values as ints.
#ifndef GRE_AMD_STATUS_HPP
#define GRE_AMD_STATUS_HPP
#include <string>

enum GRE_Status {
  GRE_SUCCESS = 0,
  GRE_FAILURE,
  GRE_NOBUFFER,
  GRE_NULL,
  GRE_FINISHED,
  GRE_UNFINISHED,
  GRE_INITIALIZED,
  GRE_UNINITIALIZED,
  GRE_NotImplemented
};

class GRE_Status_Parser {
 public:
  std::string parse(GRE_Status status) {
    switch (status) {
      case GRE_SUCCESS: return "sucess";
      case GRE_FAILURE: return "failre";
      case GRE_NOBUFFER: return "no_buffer";
      case GRE_NULL: return "nothing";
      case GRE_FINISHED: return "finished";
      case GRE_UNFINISHED: return "unfinished";
      case GRE_INITIALIZED: return "initialized";
      case GRE_UNINITIALIZED: return "uninitialized";
      case GRE_NotImplemented: return "not_implemented";
    }
    return "unknown";
  }
};
#endif
