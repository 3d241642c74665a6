// perm_index.h - the permutation stream of perm_stream.h, consumed with the
// help of per-block presence bitsets computed ahead on worker threads.
//
// One permutation's consumption (main.cu:541-558) is: n discarded draws, the
// draws up to the n1-th distinct residue (class 1), then the draws up to the
// first appearance of every other residue.  Scanned draw by draw that is
// about n * H(n) steps (4,200 for n = 600), and the stream is one sequential
// chain.  Here the chain is cut into superblocks whose starting states are
// jumped to directly (the recurrence is linear over Z/2^32, so L steps are
// one 34x34 matrix power), and worker threads produce, for each superblock,
// the raw values, their residues mod n, and for every block of kB draws the
// set of residues present in it.  The scan then moves a block at a time
// wherever the block cannot hold the draw it is looking for:
//   class 1: a block whose new residues keep the distinct count below n1 is
//            taken whole;
//   the rest: a block lacking at least one still-missing residue only removes
//            the residues it holds;
// and scans draw by draw only the partial blocks at the two ends of each
// phase.  The results (class-2 bits, stream position) are exactly those of
// PermStream::class2_bits.
#pragma once
#include <stdint.h>

#include <atomic>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "perm_stream.h"

class PermIndex {
 public:
  static const int kB = 128;       // draws per presence block
  static const int kL = 1 << 15;   // draws per superblock (a multiple of kB)
  static const int kNB = kL / kB;  // blocks per superblock
  static const int kH = 34;        // values of history (PermStream::State)

  // Starts at the stream state `st` (the position of the next draw); every
  // permutation has n samples of which n1 are class 1.
  PermIndex(const PermStream::State &st, int n, int n1, int workers) : n_(n), n1_(n1) {
    W_ = (n + 63) / 64;
    c1_.assign(W_, 0);
    miss_.assign(W_, 0);
    K_ = workers < 1 ? 1 : workers;
    R_ = 2 * K_;
    sh_ = 0;
    while ((2 << sh_) < n) sh_++;  // 2^sh < n <= 2^(sh+1)
    m_ = n > 1 ? ((uint64_t)1 << (32 + sh_)) / (uint32_t)n + 1 : 0;
    slots_.reset(new Slot[R_]);
    for (int i = 0; i < R_; i++) {
      slots_[i].raw.resize(kH + kL);
      slots_[i].res.resize(kL);
      slots_[i].pres.resize((size_t)kNB * W_);
    }
    // worker w starts at superblock w and strides by K superblocks
    const Mat &P = jump_superblock();
    const Mat Q = mat_pow(P, K_);
    std::vector<Vec> start(K_);
    for (int i = 0; i < kH; i++) start[0].v[i] = st.h[i];
    for (int w = 1; w < K_; w++) start[w] = mat_vec(P, start[w - 1]);
    for (int w = 0; w < K_; w++) pool_.emplace_back([this, w, Q, s0 = start[w]]() { work(w, Q, s0); });
  }
  ~PermIndex() {
    quit_.store(true);
    for (auto &t : pool_) t.join();
  }

  // One permutation: c2 (when given) receives the class-2 bits of samples
  // 0..n-1, the position moves past the permutation's draws.
  void class2_bits(uint32_t *c2) {
    const int n = n_, n1 = n1_, W = W_;
    uint64_t *c1 = c1_.data(), *miss = miss_.data();
    for (int w = 0; w < W; w++) c1[w] = 0;
    long long p = g_ + n;  // discarded draws
    int cnt = 0;
    // class 1: the first n1 distinct residues
    while (cnt < n1) {
      const Slot &s = get(p / kL);
      const int off = (int)(p % kL), b = off / kB, bend = (b + 1) * kB;
      ahead(s, b);
      if (off % kB == 0) {
        const uint64_t *pr = &s.pres[(size_t)b * W];
        int nw = 0;
        for (int w = 0; w < W; w++) nw += __builtin_popcountll(pr[w] & ~c1[w]);
        if (cnt + nw < n1) {
          for (int w = 0; w < W; w++) c1[w] |= pr[w];
          cnt += nw;
          p += kB;
          continue;
        }
      }
      const uint16_t *r = s.res.data();
      int i = off;
      for (; i < bend && cnt < n1; i++) {
        const uint32_t v = r[i];
        const uint64_t bit = 1ull << (v & 63);
        cnt += (c1[v >> 6] & bit) == 0;
        c1[v >> 6] |= bit;
      }
      p += i - off;
    }
    // the rest: until every other residue has appeared
    int left = n - cnt;
    for (int w = 0; w < W; w++) miss[w] = ~c1[w];
    if (n & 63) miss[W - 1] &= (1ull << (n & 63)) - 1;
    while (left > 0) {
      const Slot &s = get(p / kL);
      const int off = (int)(p % kL), b = off / kB, bend = (b + 1) * kB;
      ahead(s, b);
      if (off % kB == 0) {
        const uint64_t *pr = &s.pres[(size_t)b * W];
        uint64_t lack = 0;
        for (int w = 0; w < W; w++) lack |= miss[w] & ~pr[w];
        if (lack) {  // this block cannot complete the permutation
          for (int w = 0; w < W; w++) {
            left -= __builtin_popcountll(miss[w] & pr[w]);
            miss[w] &= ~pr[w];
          }
          p += kB;
          continue;
        }
      }
      const uint16_t *r = s.res.data();
      int i = off;
      for (; i < bend && left > 0; i++) {
        const uint32_t v = r[i];
        const uint64_t bit = 1ull << (v & 63);
        left -= (miss[v >> 6] & bit) != 0;
        miss[v >> 6] &= ~bit;
      }
      p += i - off;
    }
    g_ = p;
    if (c2) {
      const int words = (n + 31) / 32;
      for (int i = 0; i < words; i++) c2[i] = ~(uint32_t)(c1[i >> 1] >> (32 * (i & 1)));
      if (n & 31) c2[words - 1] &= (1u << (n & 31)) - 1u;
    }
  }

  // The stream state at the current position (the 34 values before it).
  PermStream::State save() {
    const Slot &s = get(g_ / kL);
    const int j = (int)(g_ % kL);
    PermStream::State st;
    for (int i = 0; i < kH; i++) st.h[i] = s.raw[j + i];
    return st;
  }

 private:
  struct Mat {
    uint32_t m[kH][kH];
  };
  struct Vec {
    uint32_t v[kH];
  };
  struct Slot {
    std::atomic<long long> k{-1};  // superblock held (complete) or -1
    std::vector<uint32_t> raw;     // kH values before the superblock, then its kL values
    std::vector<uint16_t> res;     // (raw >> 1) mod n
    std::vector<uint64_t> pres;    // per block: residues present
  };

  // x_{i+1} = M x_i with x_i = (r[i-34], .., r[i-1]) and r[i] = r[i-31] + r[i-3],
  // all arithmetic mod 2^32
  static Mat mat_mul(const Mat &a, const Mat &b) {
    Mat c;
    for (int i = 0; i < kH; i++)
      for (int j = 0; j < kH; j++) {
        uint32_t acc = 0;
        for (int k = 0; k < kH; k++) acc += a.m[i][k] * b.m[k][j];
        c.m[i][j] = acc;
      }
    return c;
  }
  static Mat mat_pow(Mat base, long long e) {
    Mat r;
    for (int i = 0; i < kH; i++)
      for (int j = 0; j < kH; j++) r.m[i][j] = i == j;
    while (e > 0) {
      if (e & 1) r = mat_mul(r, base);
      e >>= 1;
      if (e) base = mat_mul(base, base);
    }
    return r;
  }
  static Vec mat_vec(const Mat &a, const Vec &x) {
    Vec y;
    for (int i = 0; i < kH; i++) {
      uint32_t acc = 0;
      for (int k = 0; k < kH; k++) acc += a.m[i][k] * x.v[k];
      y.v[i] = acc;
    }
    return y;
  }
  // M^kL: one superblock ahead
  static const Mat &jump_superblock() {
    static const Mat P = []() {
      Mat M;
      for (int i = 0; i < kH; i++)
        for (int j = 0; j < kH; j++) M.m[i][j] = 0;
      for (int i = 0; i + 1 < kH; i++) M.m[i][i + 1] = 1;
      M.m[kH - 1][kH - 31] = 1;  // r[i-31]
      M.m[kH - 1][kH - 3] = 1;   // r[i-3]
      return mat_pow(M, kL);
    }();
    return P;
  }

  // The scan reads what other cores wrote: ask for the presence words a few
  // blocks ahead and the residues of the next block before they are needed.
  inline void ahead(const Slot &s, int b) const {
    if (b + 8 < kNB) {
      const char *q = (const char *)&s.pres[(size_t)(b + 8) * W_];
      __builtin_prefetch(q, 0, 3);
      __builtin_prefetch(q + 8 * W_ - 1, 0, 3);
    }
    if (b + 1 < kNB) {
      const char *q = (const char *)&s.res[(size_t)(b + 1) * kB];
      for (int i = 0; i < kB * 2; i += 64) __builtin_prefetch(q + i, 0, 3);
    }
  }

  // superblock k, once a worker has filled it; the superblocks before k are
  // then no longer read (the scan only moves forward, and the state saved at
  // the end of a permutation lies in the superblock of its position)
  const Slot &get(long long k) {
    if (k > release_.load(std::memory_order_relaxed)) release_.store(k, std::memory_order_release);
    Slot &s = slots_[k % R_];
    int spins = 0;
    while (s.k.load(std::memory_order_acquire) != k) {
      if (++spins > 64) std::this_thread::yield();
    }
    return s;
  }

  void work(int w, Mat Q, Vec st) {
    for (long long k = w;; k += K_) {
      int spins = 0;
      while (k - R_ >= release_.load(std::memory_order_acquire)) {
        if (quit_.load(std::memory_order_relaxed)) return;
        if (++spins > 64) std::this_thread::yield();
      }
      if (quit_.load(std::memory_order_relaxed)) return;
      Slot &s = slots_[k % R_];
      fill(s, st);
      s.k.store(k, std::memory_order_release);
      st = mat_vec(Q, st);
    }
  }


  void fill(Slot &s, const Vec &st) {
    uint32_t *b = s.raw.data();
    for (int i = 0; i < kH; i++) b[i] = st.v[i];
    uint32_t a0 = b[kH - 3], a1 = b[kH - 2], a2 = b[kH - 1];
    int i = kH;
    for (; i + 3 <= kH + kL; i += 3) {
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
    for (; i < kH + kL; i++) b[i] = b[i - 31] + b[i - 3];
    residues(b + kH, s.res.data());
    uint64_t *pr = s.pres.data();
    std::memset(pr, 0, sizeof(uint64_t) * (size_t)kNB * W_);
    const uint16_t *r = s.res.data();
    for (int blk = 0; blk < kNB; blk++) {
      uint64_t *q = pr + (size_t)blk * W_;
      for (int j = blk * kB; j < (blk + 1) * kB; j++) q[r[j] >> 6] |= 1ull << (r[j] & 63);
    }
  }
  inline uint16_t mod_n(uint32_t x) const {
    if (n_ <= 1) return 0;
    const uint32_t q = (uint32_t)(((uint64_t)x * m_) >> (32 + sh_));
    return (uint16_t)(x - q * (uint32_t)n_);
  }
  void residues(const uint32_t *raw, uint16_t *res) const {
#if defined(__x86_64__)
    if (n_ > 1 && avx512()) {
      residues_avx512(raw, res);
      return;
    }
#endif
    for (int i = 0; i < kL; i++) res[i] = mod_n(raw[i] >> 1);
  }
#if defined(__x86_64__)
  static bool avx512() {
    static const bool ok = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw");
    return ok;
  }
  // the quotient (x m) >> (32 + s) in 64-bit lanes, as in perm_stream.h
  __attribute__((target("avx512f,avx512bw"))) void residues_avx512(const uint32_t *raw, uint16_t *res) const {
    const __m512i vm = _mm512_set1_epi64((long long)m_);
    const __m512i vn = _mm512_set1_epi32(n_);
    const __m128i shv = _mm_cvtsi32_si128(32 + sh_);
    const __m512i lo32 = _mm512_set1_epi64(0xFFFFFFFFLL);
    for (int i = 0; i < kL; i += 16) {
      const __m512i x = _mm512_srli_epi32(_mm512_loadu_si512((const void *)(raw + i)), 1);
      const __m512i qe = _mm512_srl_epi64(_mm512_mul_epu32(x, vm), shv);
      const __m512i qo = _mm512_srl_epi64(_mm512_mul_epu32(_mm512_srli_epi64(x, 32), vm), shv);
      const __m512i q = _mm512_or_si512(_mm512_and_si512(qe, lo32), _mm512_slli_epi64(qo, 32));
      const __m512i v = _mm512_sub_epi32(x, _mm512_mullo_epi32(q, vn));
      _mm256_storeu_si256((__m256i *)(res + i), _mm512_cvtepi32_epi16(v));
    }
  }
#endif

  int n_, n1_, W_, K_, R_;
  std::vector<uint64_t> c1_, miss_;  // class-1 residues and the residues still missing (W_ words each)
  int sh_ = 0;
  uint64_t m_ = 0;
  long long g_ = 0;  // draws consumed since the starting state
  std::unique_ptr<Slot[]> slots_;
  std::atomic<long long> release_{0};  // superblocks below this one are no longer read
  std::atomic<bool> quit_{false};
  std::vector<std::thread> pool_;
};
