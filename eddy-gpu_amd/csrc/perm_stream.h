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

#include <algorithm>
#include <vector>

class PermStream {
 public:
  PermStream() { reset(1); }
  void reset(uint32_t seed) {
    int32_t s = (int32_t)(seed == 0 ? 1 : seed);
    uint32_t r[34];
    r[0] = (uint32_t)s;
    int32_t w = s;
    for (int i = 1; i < 31; i++) {
      int32_t hi = w / 127773, lo = w % 127773;
      w = 16807 * lo - 2836 * hi;
      if (w < 0) w += 2147483647;
      r[i] = (uint32_t)w;
    }
    for (int i = 31; i < 34; i++) r[i] = r[i - 31];
    // buf_ holds r[base_ - 34 .. base_ + kBlock): the first 34 entries are
    // the recurrence's history, the rest the next block of values
    for (int i = 0; i < 34; i++) buf_[i] = r[i];
    pos_ = kHist;
    end_ = kHist;  // empty block: the next draw refills
    for (int k = 34; k < 344; k++) next_raw();
  }
  // next rand() value in [0, 2^31)
  inline uint32_t next() { return next_raw() >> 1; }

  // The recurrence's whole state: the 34 values before the next one.
  struct State {
    uint32_t h[34];
  };
  State save() const {
    State st;
    for (int i = 0; i < kHist; i++) st.h[i] = buf_[pos_ - kHist + i];
    return st;
  }
  void restore(const State &st) {
    for (int i = 0; i < kHist; i++) buf_[i] = st.h[i];
    pos_ = end_ = kHist;  // empty block: the next draw refills from the history
  }

  // One permutation's consumption (main.cu:541-558): `n` discarded draws,
  // then rejection sampling of a permutation of [0, n).  `rand() % n` is
  // computed with an exact multiply-shift reduction (Lemire, "Faster
  // remainder by direct computation", 2019), which equals % for every 32-bit
  // input.  Only what the class labels need is kept: the first n1 accepted
  // values are class 1 (main.cu:561-570), so the scan records those, then only
  // counts distinct values until all n are seen (where the next permutation's
  // draws begin).  c2 (when given) receives the class-2 membership bits of
  // samples 0..n-1.
  void class2_bits(int n, int n1, uint32_t *c2) {
    if ((int)stamp_.size() < n) stamp_.assign(n, 0);
    if (++gen_ == 0) {
      std::fill(stamp_.begin(), stamp_.end(), 0);
      gen_ = 1;
    }
    const int words = (n + 31) / 32;
    if (c2)
      for (int i = 0; i < words; i++) c2[i] = 0u;
    const uint64_t M = UINT64_C(0xFFFFFFFFFFFFFFFF) / (uint32_t)n + 1;
    skip(n);
    uint32_t *stamp = stamp_.data();
    const uint32_t gen = gen_;
    int k = 0;
    // class 1: the first n1 distinct values (marked in c2, inverted below)
    while (k < n1) {
      if (pos_ == end_) refill();
      const int avail = end_ - pos_;
      const uint32_t *src = buf_ + pos_;
      int i = 0;
      for (; i < avail && k < n1; i++) {
        const uint32_t x = src[i] >> 1;
        const uint32_t v = (uint32_t)(((__uint128_t)(M * x) * (uint32_t)n) >> 64);
        const uint32_t fresh = stamp[v] != gen;
        stamp[v] = gen;
        if (c2) c2[v >> 5] |= fresh << (v & 31);
        k += (int)fresh;
      }
      pos_ += i;
    }
    // the rest: count until every value has appeared
    while (k < n) {
      if (pos_ == end_) refill();
      const int avail = end_ - pos_;
      const uint32_t *src = buf_ + pos_;
      int i = 0;
      for (; i < avail && k < n; i++) {
        const uint32_t x = src[i] >> 1;
        const uint32_t v = (uint32_t)(((__uint128_t)(M * x) * (uint32_t)n) >> 64);
        k += (int)(stamp[v] != gen);
        stamp[v] = gen;
      }
      pos_ += i;
    }
    if (c2) {
      for (int i = 0; i < words; i++) c2[i] = ~c2[i];
      if (n & 31) c2[words - 1] &= (1u << (n & 31)) - 1u;
    }
  }

 private:
  static const int kHist = 34;
  static const int kBlock = 8190;  // a multiple of 3 (the refill loop's step)
  // next block: r[i] = r[i-31] + r[i-3] over a linear array (3-way independent)
  void refill() {
    for (int i = 0; i < kHist; i++) buf_[i] = buf_[end_ - kHist + i];
    uint32_t *b = buf_;
    // the three most recent values ride in registers: three independent
    // add chains instead of a store-to-load round trip per value
    uint32_t a0 = b[kHist - 3], a1 = b[kHist - 2], a2 = b[kHist - 1];
    for (int i = kHist; i < kHist + kBlock; i += 3) {
      const uint32_t v0 = b[i - 31] + a0;
      const uint32_t v1 = b[i - 30] + a1;
      const uint32_t v2 = b[i - 29] + a2;
      b[i] = v0;
      b[i + 1] = v1;
      b[i + 2] = v2;
      a0 = v0;
      a1 = v1;
      a2 = v2;
    }
    pos_ = kHist;
    end_ = kHist + kBlock;
  }
  inline uint32_t next_raw() {
    if (pos_ == end_) refill();
    return buf_[pos_++];
  }
  void skip(int n) {
    while (n > 0) {
      if (pos_ == end_) refill();
      int m = std::min(n, end_ - pos_);
      pos_ += m;
      n -= m;
    }
  }
  uint32_t buf_[kHist + kBlock];
  int pos_ = 0, end_ = 0;
  std::vector<uint32_t> stamp_;
  uint32_t gen_ = 0;
};
