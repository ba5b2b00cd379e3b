// Small helpers GRE programs and drivers call (include/Util/Util.hpp:90-157 in
// the GRE tree): align(), toString(), interpretWord(), loadFromFile().
#ifndef GRE_AMD_UTIL_HPP
#define GRE_AMD_UTIL_HPP
#include <stdint.h>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "Types.hpp"

namespace GRE {
namespace Util {

template <typename T>
std::string toString(const T& val) {
  std::ostringstream ss;
  ss << val;
  return ss.str();
}

// Ids of the set bits of a 64-bit word, offset by base; returns the count.
template <typename T>
int interpretWord(T* buf, uint64_t bits, const int base) {
  int n = 0;
  while (bits) {
    const int b = __builtin_ctzll(bits);
    buf[n++] = (T)(base + b);
    bits &= bits - 1;
  }
  return n;
}

// Rounds value up to a multiple of size (64 by default): the padding GRE puts
// after the master vertices in every per-vertex array.
template <typename T>
T align(T value, const int size = 64) {
  const T r = value % size;
  return r == 0 ? value : value + (size - r);
}
template <typename T>
T uppAlign(T value) { return value + (64 - value % 64); }
template <typename T>
T downAlign(T value) { return value - value % 64; }

// Reads whitespace-separated values into array; stops after n values when n>0.
// Returns the number of extraction attempts, counting the final one that hits
// end-of-file, which is how GRE's drivers index a ".roots" file: element 0 is
// the count header and element 1 the first root.
template <typename T>
size_t loadFromFile(const std::string path, std::vector<T>& array, const size_t n = 0) {
  std::ifstream in(path.c_str());
  if (!in) {
    std::cerr << "Err: fail to read data from" << path << std::endl;
    return 0;
  }
  size_t count = 0;
  while (in.good() && !in.eof()) {
    T elem = T();
    in >> elem;
    array.push_back(elem);
    ++count;
    if (n != 0 && count == n) break;
  }
  return count;
}

}  // namespace Util
}  // namespace GRE
#endif
