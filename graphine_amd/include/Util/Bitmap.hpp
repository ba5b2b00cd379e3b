// Host bitmap with GRE's interface (include/Util/Bitmap.hpp:12-130 in the GRE
// tree).  The device frontiers are separate bitmaps in HBM; this class serves
// host code that uses GRE's Bitmap directly.
#ifndef GRE_AMD_BITMAP_HPP
#define GRE_AMD_BITMAP_HPP
#include <stdint.h>
#include <cstdlib>
#include <cstring>

namespace GRE {
namespace Util {

class Bitmap {
 public:
  typedef uint64_t* wordIterator;
  Bitmap() : size_(0), bits_(NULL) {}
  explicit Bitmap(int64_t nBits) : size_(0), bits_(NULL) { resize(nBits); }
  ~Bitmap() { destroy(); }
  void destroy() {
    std::free(bits_);
    bits_ = NULL;
  }
  bool resize(int64_t nBits) {
    if (nBits > size_) {
      std::free(bits_);
      size_ = nBits;
      bits_ = static_cast<uint64_t*>(std::calloc(nWords(), sizeof(uint64_t)));
    }
    return bits_ != NULL;
  }
  void set(int64_t i) { __sync_fetch_and_or(&bits_[i >> 6], 1ULL << (i & 63)); }
  void unset(int64_t i) { __sync_fetch_and_and(&bits_[i >> 6], ~(1ULL << (i & 63))); }
  bool test(int64_t i) const { return (bits_[i >> 6] >> (i & 63)) & 1ULL; }
  bool testAndSet(int64_t i) {
    const uint64_t m = 1ULL << (i & 63);
    if (bits_[i >> 6] & m) return true;
    return (__sync_fetch_and_or(&bits_[i >> 6], m) & m) != 0;
  }
  void set_serial(int64_t i) { bits_[i >> 6] |= 1ULL << (i & 63); }
  void unset_serial(int64_t i) { bits_[i >> 6] &= ~(1ULL << (i & 63)); }
  int64_t size() const { return size_; }
  int64_t nWords() const { return size_ / 64 + 1; }
  void clear() {
    if (bits_) std::memset(bits_, 0, sizeof(uint64_t) * nWords());
  }
  void reset() { clear(); }
  uint64_t getWord(int64_t w) const { return bits_[w]; }
  void setWord(int64_t w, uint64_t v) { bits_[w] = v; }
  wordIterator begin() { return bits_; }
  wordIterator end() { return bits_ + nWords(); }
  wordIterator getWordIterator(int64_t id) { return bits_ + (id >> 6); }

 private:
  Bitmap(const Bitmap&);
  Bitmap& operator=(const Bitmap&);
  int64_t size_;
  uint64_t* bits_;
};

template <size_t N = 64>
class FixedBitmap {
 public:
  FixedBitmap() { clear(); }
  void set(int64_t i) { __sync_fetch_and_or(&bits_[i >> 6], 1ULL << (i & 63)); }
  bool test(int64_t i) const { return (bits_[i >> 6] >> (i & 63)) & 1ULL; }
  void set_serial(int64_t i) { bits_[i >> 6] |= 1ULL << (i & 63); }
  void clear() { std::memset(bits_, 0, sizeof(bits_)); }
  int64_t size() const { return N; }

 private:
  uint64_t bits_[N / 64];
};

}  // namespace Util
}  // namespace GRE
#endif
