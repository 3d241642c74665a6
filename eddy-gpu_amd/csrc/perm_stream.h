// perm_stream.h - the reference's permutation stream (main.cu:538-573).
//
// The reference draws its class-label permutations from glibc rand(), never
// seeded, continuing across gene sets.  glibc's rand() is the additive
// feedback generator of random_r's TYPE_3 state (degree 31, separation 3):
//   r[0] = seed; r[i] = 16807 * r[i-1] mod (2^31 - 1)   (i = 1..30, Schrage)
//   r[i] = r[i-31]                                       (i = 31..33)
//   r[i] = r[i-31] + r[i-3] mod 2^32                     (i >= 34)
//   k-th output = r[k + 344] >> 1
// This class restates that recurrence (checked against the C library's own
// rand() in the tests) so the engine can replay the stream without a lock and
// without touching the process-wide rand() state.
#pragma once
#include <stdint.h>

#include <vector>

class PermStream {
 public:
  PermStream() { reset(1); }
  void reset(uint32_t seed) {
    int32_t s = (int32_t)(seed == 0 ? 1 : seed);
    r_[0] = (uint32_t)s;
    int32_t w = s;
    for (int i = 1; i < 31; i++) {
      int32_t hi = w / 127773, lo = w % 127773;
      w = 16807 * lo - 2836 * hi;
      if (w < 0) w += 2147483647;
      r_[i] = (uint32_t)w;
    }
    for (int i = 31; i < 34; i++) r_[i] = r_[i - 31];
    i_ = 34;
    for (int k = 34; k < 344; k++) next_raw();
  }
  // next rand() value in [0, 2^31)
  inline uint32_t next() { return next_raw() >> 1; }

  // One permutation's consumption (main.cu:541-558): `n` discarded draws,
  // then rejection sampling of a permutation of [0, n).  order receives the
  // accepted values in acceptance order.
  void permutation(int n, std::vector<int> &order) {
    order.resize(n);
    if ((int)stamp_.size() < n) stamp_.assign(n, 0);
    if (++gen_ == 0) {
      std::fill(stamp_.begin(), stamp_.end(), 0);
      gen_ = 1;
    }
    for (int c = 0; c < n; c++) next_raw();
    int k = 0;
    while (k < n) {
      uint32_t v = next() % (uint32_t)n;
      if (stamp_[v] != gen_) {
        stamp_[v] = gen_;
        order[k++] = (int)v;
      }
    }
  }

 private:
  inline uint32_t next_raw() {
    uint32_t v = r_[(i_ - 31) & 63] + r_[(i_ - 3) & 63];
    r_[i_ & 63] = v;
    i_++;
    return v;
  }
  uint32_t r_[64];
  uint64_t i_ = 0;
  std::vector<uint32_t> stamp_;
  uint32_t gen_ = 0;
};
