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
#if defined(__x86_64__)
#include <immintrin.h>
#endif

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
  // then rejection sampling of a permutation of [0, n) from rand() % n.
  // Only what the class labels need is kept: the first n1 accepted values
  // are class 1 (main.cu:561-570), so the scan records that set, then only
  // waits until every other value has appeared (where the next permutation's
  // draws begin).  c2 (when given) receives the class-2 membership bits of
  // samples 0..n-1.
  //
  // The residues rand() % n of each block are computed once per block (in
  // vector form where the CPU has AVX-512) as x - n * ((x * m) >> (32 + s))
  // with 2^s < n <= 2^(s+1) and m = floor(2^(32+s) / n) + 1 < 2^32: with
  // e = m n - 2^(32+s) in (0, n] and x < 2^31, e x < n 2^31 < 2^(32+s), so
  // the quotient is exact (Granlund & Montgomery, "Division by invariant
  // integers using multiplication", 1994).  Once at most kTail values are
  // still unseen, the scan only compares residues with those values.
  void class2_bits(int n, int n1, uint32_t *c2) {
    set_modulus(n);
    if ((int)stamp_.size() < n) stamp_.assign(n, 0);
    if (++gen_ == 0) {
      std::fill(stamp_.begin(), stamp_.end(), 0);
      gen_ = 1;
    }
    const int words = (n + 31) / 32;
    if (c2)
      for (int i = 0; i < words; i++) c2[i] = 0u;
    skip(n);
    uint32_t *stamp = stamp_.data();
    const uint32_t gen = gen_;
    int k = 0;
    // class 1: the first n1 distinct values (marked in c2, inverted below)
    while (k < n1) {
      if (pos_ == end_) refill();
      const int avail = end_ - pos_;
      const uint16_t *src = res_ + pos_;
      int i = 0;
      for (; i < avail && k < n1; i++) {
        const uint32_t v = src[i];
        const uint32_t fresh = stamp[v] != gen;
        stamp[v] = gen;
        if (c2) c2[v >> 5] |= fresh << (v & 31);
        k += (int)fresh;
      }
      pos_ += i;
    }
    // the rest: count until every value has appeared
    const int tail = std::min(n, kTail);
    while (k < n - tail) {
      if (pos_ == end_) refill();
      const int avail = end_ - pos_;
      const uint16_t *src = res_ + pos_;
      int i = 0;
      for (; i < avail && k < n - tail; i++) {
        const uint32_t v = src[i];
        k += (int)(stamp[v] != gen);
        stamp[v] = gen;
      }
      pos_ += i;
    }
    if (k < n) {
      uint16_t tg[kTail];
      int left = 0;
      for (int v = 0; v < n; v++)
        if (stamp[v] != gen) tg[left++] = (uint16_t)v;
      while (left > 0) {
        if (pos_ == end_) refill();
        pos_ = scan_targets(pos_, tg, &left);
      }
    }
    if (c2) {
      for (int i = 0; i < words; i++) c2[i] = ~c2[i];
      if (n & 31) c2[words - 1] &= (1u << (n & 31)) - 1u;
    }
  }

 private:
  static const int kHist = 34;
  static const int kBlock = 8190;  // a multiple of 3 (the refill loop's step)
  static const int kTail = 8;
  // modulus of the residues in res_; a change recomputes the current block's
  void set_modulus(int n) {
    if (n == vn_) return;
    vn_ = n;
    sh_ = 0;
    while ((2 << sh_) < n) sh_++;  // 2^sh < n <= 2^(sh+1)
    m_ = n > 1 ? ((uint64_t)1 << (32 + sh_)) / (uint32_t)n + 1 : 0;
    residues(pos_, end_);
  }
  void residues(int b, int e) {
    if (vn_ <= 0) return;
#if defined(__x86_64__)
    if (has_avx512()) {
      residues_avx512(b, e);
      return;
    }
#endif
    for (int i = b; i < e; i++) res_[i] = mod_n(buf_[i] >> 1);
  }
  inline uint16_t mod_n(uint32_t x) const {
    if (vn_ == 1) return 0;
    const uint32_t q = (uint32_t)(((uint64_t)x * m_) >> (32 + sh_));
    return (uint16_t)(x - q * (uint32_t)vn_);
  }
  // first draws at or after i that complete the targets; returns the
  // position after the last target's first appearance, or end_
  int scan_targets(int i, uint16_t *tg, int *left) {
#if defined(__x86_64__)
    if (has_avx512()) i = scan_targets_avx512(i, tg, left);
#endif
    for (; i < end_ && *left > 0; i++) take_target(res_[i], tg, left);
    return i;
  }
  static inline void take_target(uint16_t v, uint16_t *tg, int *left) {
    for (int j = 0; j < *left; j++)
      if (tg[j] == v) {
        tg[j] = tg[--*left];
        return;
      }
  }
#if defined(__x86_64__)
  static bool has_avx512() {
    static const bool ok = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw");
    return ok;
  }
  __attribute__((target("avx512f,avx512bw"))) void residues_avx512(int b, int e) {
    if (vn_ == 1) {
      for (int i = b; i < e; i++) res_[i] = 0;
      return;
    }
    const __m512i vm = _mm512_set1_epi64((long long)m_);
    const __m512i vn = _mm512_set1_epi32(vn_);
    const __m128i shv = _mm_cvtsi32_si128(32 + sh_);
    const __m512i lo32 = _mm512_set1_epi64(0xFFFFFFFFLL);
    int i = b;
    for (; i + 16 <= e; i += 16) {
      const __m512i x = _mm512_srli_epi32(_mm512_loadu_si512((const void *)(buf_ + i)), 1);
      const __m512i qe = _mm512_srl_epi64(_mm512_mul_epu32(x, vm), shv);
      const __m512i qo = _mm512_srl_epi64(_mm512_mul_epu32(_mm512_srli_epi64(x, 32), vm), shv);
      const __m512i q = _mm512_or_si512(_mm512_and_si512(qe, lo32), _mm512_slli_epi64(qo, 32));
      const __m512i v = _mm512_sub_epi32(x, _mm512_mullo_epi32(q, vn));
      _mm256_storeu_si256((__m256i *)(res_ + i), _mm512_cvtepi32_epi16(v));
    }
    for (; i < e; i++) res_[i] = mod_n(buf_[i] >> 1);
  }
  // 32 residues per step compared with every remaining target; a step with
  // a match is resolved lane by lane in order
  __attribute__((target("avx512f,avx512bw"))) int scan_targets_avx512(int i, uint16_t *tg, int *left) {
    while (*left > 0 && i + 32 <= end_) {
      const __m512i v = _mm512_loadu_si512((const void *)(res_ + i));
      __mmask32 hit = 0;
      for (int j = 0; j < *left; j++) hit |= _mm512_cmpeq_epi16_mask(v, _mm512_set1_epi16((short)tg[j]));
      if (hit == 0) {
        i += 32;
        continue;
      }
      for (int l = 0; l < 32; l++) {
        if (!((hit >> l) & 1)) continue;
        take_target(res_[i + l], tg, left);
        if (*left == 0) return i + l + 1;
      }
      i += 32;
    }
    return i;
  }
#endif
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
    residues(kHist, kHist + kBlock);
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
  uint16_t res_[kHist + kBlock];  // rand() % vn_ of buf_[i] (valid from kHist)
  int pos_ = 0, end_ = 0;
  int vn_ = 0, sh_ = 0;
  uint64_t m_ = 0;
  std::vector<uint32_t> stamp_;
  uint32_t gen_ = 0;
};
