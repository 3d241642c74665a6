// perm_chain.h - the permutation stream of perm_stream.h, consumed in parallel.
//
// One permutation's consumption (main.cu:541-558) is: n discarded draws, the
// draws up to the n1-th distinct residue rand() % n (class 1), then the draws
// up to the first appearance of every other residue.  Two facts split it:
//   - where a permutation ends depends only on where it starts: the end is
//     the first position at which every residue has appeared since the
//     discarded draws (the class split does not move it);
//   - the class split depends only on the stream state at the first kept
//     draw (34 values), and needs only the draws up to the n1-th distinct
//     residue.
// So the stream's permutations form a chain of start positions p -> f(p), and
// the split of each is recomputed from one state (`class2_from_state`), by
// any thread, in any order.
//
// The chain is walked in segments on several threads at once.  A segment's
// walk starts at a guess (a few standard deviations before where its first
// permutation is expected to start, judged from the last validated start when
// the segment is taken up), generates its own stretch of the stream
// (the recurrence is linear over Z/2^32, so any position is one product of
// 34x34 matrix powers away from the start) and records the start positions
// it passes and the state at each one's first kept draw.  Walks from
// different positions merge: f maps many positions to one, and a walk from
// an arbitrary position lands on the true chain within a few permutations.
// A merger thread stitches the segments in order at the first position the
// validated chain shares with the next segment's walk (from there on the two
// are one chain); a segment that never meets it is replaced by walking on
// from the validated end.  The validated entries are exactly the
// permutations PermStream::class2_bits draws one after another.
//
// Inside a walk, the stream is cut into superblocks of kL draws, each made on
// the walking thread (its values 16 at a time, their residues, a 34-value
// checkpoint per block of kB draws, the set of residues present per block);
// a permutation takes whole blocks while a block lacks one of the residues
// still missing, and ends inside the first block that holds them all.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#if defined(__linux__)
#include <sys/resource.h>
#include <sys/syscall.h>
#include <unistd.h>
#endif

#include "perm_stream.h"

class PermChain {
 public:
  static const int kB = 128;       // draws per presence block
  static const int kL = 1 << 15;   // draws per superblock (a multiple of kB)
  static const int kNB = kL / kB;  // blocks per superblock
  static const int kH = 34;        // values of history (PermStream::State)
  static const int kC = 256;       // draws generated at a time by class2_from_state

  // Per-thread working memory of class2_from_state.
  struct Scratch {
    std::vector<uint32_t> raw;  // kH values of history, then kC values
    std::vector<uint16_t> res;  // residues of the kC values
    std::vector<uint64_t> c1;   // class-1 residues so far
  };

  // The chain of `count` permutations from the stream state `base` (the
  // position of the next draw, position 0), every permutation of n samples
  // of which n1 are class 1, walked by `threads` threads.
  PermChain(const PermStream::State &base, int n, int n1, long long count, int threads, int nice = 0)
      : g_(n, n1), count_(count < 0 ? 0 : count), nice_(nice) {
    for (int i = 0; i < kH; i++) base_.v[i] = base.h[i];
    pos_.assign(count_ + 1, 0);
    kept_.resize(count_ + 1);
    plan();
    segs_.reset(new Seg[nseg_]);
    const int T = std::max(1, threads);
    for (int t = 0; t < T; t++) pool_.emplace_back([this]() { seg_loop(); });
    merger_ = std::thread([this]() { merge_loop(); });
  }
  ~PermChain() {
    quit_.store(true);
    {
      std::lock_guard<std::mutex> g(mu_);
      cv_.notify_all();
      vcv_.notify_all();
    }
    for (auto &t : pool_) t.join();
    if (merger_.joinable()) merger_.join();
  }

  // entries 0 .. validated() - 1 are final (start position and kept state)
  long long validated() const { return valid_.load(std::memory_order_acquire); }
  // sleeps until entry i is final; false if the chain is stopping first
  bool wait(long long i) {
    if (validated() > i) return true;
    const auto w0 = std::chrono::steady_clock::now();
    std::unique_lock<std::mutex> g(mu_);
    vcv_.wait(g, [&]() { return validated() > i || quit_.load(std::memory_order_relaxed); });
    wait_us_ += std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - w0).count();
    return validated() > i;
  }
  // the stream state at permutation i's first kept draw (i < count)
  const PermStream::State &kept(long long i) const { return kept_[i]; }
  // the stream state at permutation i's start (i <= count): the state after
  // permutations 0 .. i - 1
  PermStream::State start_state(long long i) const {
    PermStream::State st;
    state_at_position(pos_[i], st.h);
    return st;
  }
  double wait_ms() const { return 1e-3 * (double)wait_us_.load(); }
  int segments() const { return nseg_; }
  long long fallback_steps() const { return fallback_.load(); }

  // The class-2 bits (c2, (n + 31) / 32 words) of the permutation whose first
  // kept draw follows the stream state st: class 1 is the first n1 distinct
  // residues rand() % n (main.cu:541-570).  Reads nothing shared.
  void class2_from_state(const PermStream::State &st, uint32_t *c2, Scratch &sc) const {
    const int n = g_.n, n1 = g_.n1, W = g_.W;
    if ((int)sc.raw.size() < kH + kC) sc.raw.resize(kH + kC);
    if ((int)sc.res.size() < kC) sc.res.resize(kC);
    if ((int)sc.c1.size() < W) sc.c1.resize(W);
    uint32_t *b = sc.raw.data();
    uint16_t *r = sc.res.data();
    uint64_t *c1 = sc.c1.data();
    for (int w = 0; w < W; w++) c1[w] = 0;
    for (int i = 0; i < kH; i++) b[i] = st.h[i];
    int cnt = 0;
    while (cnt < n1) {
      uint32_t a0 = b[kH - 3], a1 = b[kH - 2], a2 = b[kH - 1];
      static_assert(kC % 3 == 1, "the loop below leaves one value");
      int i = kH;
      for (; i + 3 <= kH + kC; i += 3) {
        const uint32_t v0 = b[i - 31] + a0, v1 = b[i - 30] + a1, v2 = b[i - 29] + a2;
        b[i] = v0;
        b[i + 1] = v1;
        b[i + 2] = v2;
        a0 = v0;
        a1 = v1;
        a2 = v2;
      }
      b[i] = b[i - 31] + b[i - 3];
      g_.residues(b + kH, r, kC);
      for (int k = 0; k < kC && cnt < n1; k++) {
        const uint32_t v = r[k];
        const uint64_t bit = 1ull << (v & 63);
        cnt += (c1[v >> 6] & bit) == 0;
        c1[v >> 6] |= bit;
      }
      for (int j = 0; j < kH; j++) b[j] = b[kC + j];
    }
    const int words = (n + 31) / 32;
    for (int i = 0; i < words; i++) c2[i] = ~(uint32_t)(c1[i >> 1] >> (32 * (i & 1)));
    if (n & 31) c2[words - 1] &= (1u << (n & 31)) - 1u;
  }

 private:
  struct Mat {
    uint32_t m[kH][kH];
  };
  struct Vec {
    uint32_t v[kH];
  };
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
  static Vec mat_vec(const Mat &a, const Vec &x) {
    Vec y;
    for (int i = 0; i < kH; i++) {
      uint32_t acc = 0;
      for (int k = 0; k < kH; k++) acc += a.m[i][k] * x.v[k];
      y.v[i] = acc;
    }
    return y;
  }
  // x_{i+1} = M x_i with x_i = (r[i-34], .., r[i-1]) and r[i] = r[i-31] +
  // r[i-3] (mod 2^32); jumps()[e] = M^(kL 2^e): superblock k starts at
  // M^(kL k) applied to the base state
  static const int kJumps = 40;
  static const std::vector<Mat> &jumps() {
    static const std::vector<Mat> J = []() {
      Mat M;
      for (int i = 0; i < kH; i++)
        for (int j = 0; j < kH; j++) M.m[i][j] = 0;
      for (int i = 0; i + 1 < kH; i++) M.m[i][i + 1] = 1;
      M.m[kH - 1][kH - 31] = 1;
      M.m[kH - 1][kH - 3] = 1;
      Mat P = M;  // M^(2^e), up to M^kL
      for (int e = 0; (1 << e) < kL; e++) P = mat_mul(P, P);
      std::vector<Mat> v(kJumps);
      v[0] = P;
      for (int e = 1; e < kJumps; e++) v[e] = mat_mul(v[e - 1], v[e - 1]);
      return v;
    }();
    return J;
  }
  Vec superblock_start(long long k) const {
    Vec x = base_;
    const std::vector<Mat> &J = jumps();
    for (int e = 0; e < kJumps && (k >> e) != 0; e++)
      if ((k >> e) & 1) x = mat_vec(J[e], x);
    return x;
  }
  // the 34 values before absolute position p
  void state_at_position(long long p, uint32_t *h) const {
    const long long k = p / kL;
    const int o = (int)(p % kL);
    const Vec x = superblock_start(k);
    std::vector<uint32_t> b((size_t)kH + o);
    for (int i = 0; i < kH; i++) b[i] = x.v[i];
    for (int i = kH; i < kH + o; i++) b[i] = b[i - 31] + b[i - 3];
    for (int i = 0; i < kH; i++) h[i] = b[o + i];
  }

  struct Slot {
    std::vector<uint32_t> ckpt;  // per block: the kH values before it
    std::vector<uint16_t> res;   // (value >> 1) mod n
    std::vector<uint64_t> pres;  // per block: residues present
  };

  // The generator: a superblock's residues, checkpoints and presence words
  // from its starting state.
  struct Gen {
    int n, n1, W, sh;
    uint64_t m;
    Gen(int n_, int n1_) : n(n_), n1(n1_) {
      W = (n + 63) / 64;
      sh = 0;
      while ((2 << sh) < n) sh++;  // 2^sh < n <= 2^(sh+1)
      m = n > 1 ? ((uint64_t)1 << (32 + sh)) / (uint32_t)n + 1 : 0;
    }
    inline uint16_t mod_n(uint32_t x) const {
      if (n <= 1) return 0;
      const uint32_t q = (uint32_t)(((uint64_t)x * m) >> (32 + sh));
      return (uint16_t)(x - q * (uint32_t)n);
    }
    // residues of `count` values (a multiple of 16)
    void residues(const uint32_t *raw, uint16_t *res, int count) const {
#if defined(__x86_64__)
      if (n > 1 && avx512()) {
        residues_avx512(raw, res, count);
        return;
      }
#endif
      for (int i = 0; i < count; i++) res[i] = mod_n(raw[i] >> 1);
    }
    void fill(Slot &s, const Vec &st, std::vector<uint32_t> &raw) const {
      if (s.res.empty()) {
        s.ckpt.resize((size_t)kNB * kH);
        s.res.resize(kL);
        s.pres.resize((size_t)kNB * W);
      }
#if defined(__x86_64__)
      if (avx512() && n > 1) {
        values_avx512(st, s);
      } else
#endif
      {
        raw.resize(kH + kL);
        uint32_t *b = raw.data();
        for (int i = 0; i < kH; i++) b[i] = st.v[i];
        uint32_t a0 = b[kH - 3], a1 = b[kH - 2], a2 = b[kH - 1];
        int i = kH;
        for (; i + 3 <= kH + kL; i += 3) {
          const uint32_t v0 = b[i - 31] + a0, v1 = b[i - 30] + a1, v2 = b[i - 29] + a2;
          b[i] = v0;
          b[i + 1] = v1;
          b[i + 2] = v2;
          a0 = v0;
          a1 = v1;
          a2 = v2;
        }
        for (; i < kH + kL; i++) b[i] = b[i - 31] + b[i - 3];
        residues(b + kH, s.res.data(), kL);
        for (int blk = 0; blk < kNB; blk++)
          for (int j = 0; j < kH; j++) s.ckpt[(size_t)blk * kH + j] = b[blk * kB + j];
      }
      uint64_t *pr = s.pres.data();
      const uint16_t *r = s.res.data();
#if defined(__x86_64__)
      if (avx512() && W <= 32) {
        presence_avx512(r, pr);
        return;
      }
#endif
      std::memset(pr, 0, sizeof(uint64_t) * (size_t)kNB * W);
      for (int blk = 0; blk < kNB; blk++) {
        uint64_t *q = pr + (size_t)blk * W;
        for (int j = blk * kB; j < (blk + 1) * kB; j++) q[r[j] >> 6] |= 1ull << (r[j] & 63);
      }
    }
#if defined(__x86_64__)
    static bool avx512() {
      static const bool ok = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw");
      return ok;
    }
    // the quotient (x m) >> (32 + s) in 64-bit lanes, as in perm_stream.h
    __attribute__((target("avx512f,avx512bw"))) void residues_avx512(const uint32_t *raw, uint16_t *res,
                                                                    int count) const {
      const __m512i vm = _mm512_set1_epi64((long long)m);
      const __m512i vn = _mm512_set1_epi32(n);
      const __m128i shv = _mm_cvtsi32_si128(32 + sh);
      const __m512i lo32 = _mm512_set1_epi64(0xFFFFFFFFLL);
      for (int i = 0; i < count; i += 16) {
        const __m512i x = _mm512_srli_epi32(_mm512_loadu_si512((const void *)(raw + i)), 1);
        const __m512i qe = _mm512_srl_epi64(_mm512_mul_epu32(x, vm), shv);
        const __m512i qo = _mm512_srl_epi64(_mm512_mul_epu32(_mm512_srli_epi64(x, 32), vm), shv);
        const __m512i q = _mm512_or_si512(_mm512_and_si512(qe, lo32), _mm512_slli_epi64(qo, 32));
        const __m512i v = _mm512_sub_epi32(x, _mm512_mullo_epi32(q, vn));
        _mm256_storeu_si256((__m256i *)(res + i), _mm512_cvtepi32_epi16(v));
      }
    }
    // The values 16 at a time, turned into residues as they are made, with
    // the checkpoints taken from the vectors in registers.  Substituting
    // r[i-3] = r[i-6] + r[i-34] five times gives r[i] = r[i-18] + r[i-31] +
    // r[i-34] + r[i-37] + r[i-40] + r[i-43] + r[i-46]: every term lies at
    // least 18 back, so a vector of 16 consecutive values is a sum of windows
    // of the three vectors before it.
    __attribute__((target("avx512f,avx512bw"))) void values_avx512(const Vec &st, Slot &s) const {
      static_assert(kH == 34 && kL % kB == 0 && kB % 16 == 0, "layout");
      // positions -34 .. 13 (the history, then 14 values one at a time), so
      // that the vectors start at position 14 and a block's start, a
      // multiple of 128, is element 2 of a vector
      alignas(64) uint32_t b[48];
      for (int i = 0; i < kH; i++) b[i] = st.v[i];
      for (int i = kH; i < 48; i++) b[i] = b[i - 31] + b[i - 3];
      for (int i = 0; i < kH; i++) s.ckpt[i] = b[i];
      uint16_t *res = s.res.data();
      for (int i = 0; i < 14; i++) res[i] = mod_n(b[kH + i] >> 1);
      __m512i p3 = _mm512_load_si512((const void *)b), p2 = _mm512_load_si512((const void *)(b + 16));
      __m512i p1 = _mm512_load_si512((const void *)(b + 32));
      const __m512i vm = _mm512_set1_epi64((long long)m);
      const __m512i vn = _mm512_set1_epi32(n);
      const __m128i shv = _mm_cvtsi32_si128(32 + sh);
      const __m512i lo32 = _mm512_set1_epi64(0xFFFFFFFFLL);
      int q = 14;  // position of the vector's first value
      for (; q + 16 <= kL; q += 16) {
        // windows r[q-d .. q-d+15] of (p3 p2 p1) = r[q-48 .. q-1]
        const __m512i w31 = _mm512_alignr_epi32(p1, p2, 1), w34 = _mm512_alignr_epi32(p2, p3, 14);
        const __m512i w37 = _mm512_alignr_epi32(p2, p3, 11), w40 = _mm512_alignr_epi32(p2, p3, 8);
        const __m512i w43 = _mm512_alignr_epi32(p2, p3, 5), w46 = _mm512_alignr_epi32(p2, p3, 2);
        const __m512i far = _mm512_add_epi32(_mm512_add_epi32(_mm512_add_epi32(w31, w34), _mm512_add_epi32(w37, w40)),
                                             _mm512_add_epi32(w43, w46));
        const __m512i v = _mm512_add_epi32(far, _mm512_alignr_epi32(p1, p2, 14));
        const __m512i x = _mm512_srli_epi32(v, 1);
        const __m512i qe = _mm512_srl_epi64(_mm512_mul_epu32(x, vm), shv);
        const __m512i qo = _mm512_srl_epi64(_mm512_mul_epu32(_mm512_srli_epi64(x, 32), vm), shv);
        const __m512i qq = _mm512_or_si512(_mm512_and_si512(qe, lo32), _mm512_slli_epi64(qo, 32));
        _mm256_storeu_si256((__m256i *)(res + q),
                            _mm512_cvtepi32_epi16(_mm512_sub_epi32(x, _mm512_mullo_epi32(qq, vn))));
        if (((q + 2) & (kB - 1)) == 0) {
          // block (q + 2) / kB starts at element 2: the 34 values before it
          // are p2, p1 and elements 0, 1 of v
          uint32_t *c = &s.ckpt[(size_t)((q + 2) / kB) * kH];
          _mm512_storeu_si512((void *)c, p2);
          _mm512_storeu_si512((void *)(c + 16), p1);
          _mm512_mask_storeu_epi32((void *)(c + 32), (__mmask16)0x3, v);
        }
        p3 = p2;
        p2 = p1;
        p1 = v;
      }
      // the last values one at a time, from the three vectors in registers
      alignas(64) uint32_t t[48 + 16];
      _mm512_store_si512((void *)t, p3);
      _mm512_store_si512((void *)(t + 16), p2);
      _mm512_store_si512((void *)(t + 32), p1);
      for (int i = 48; q < kL; i++, q++) {
        t[i] = t[i - 31] + t[i - 3];
        res[q] = mod_n(t[i] >> 1);
      }
    }
    // Presence words of every block: a byte per residue set by plain stores,
    // then 64 bytes at a time packed to one word by their sign bits and
    // cleared.
    __attribute__((target("avx512f,avx512bw"))) void presence_avx512(const uint16_t *r, uint64_t *pr) const {
      alignas(64) uint8_t mk[32 * 64];
      const __m512i zero = _mm512_setzero_si512();
      for (int w = 0; w < W; w++) _mm512_store_si512((void *)(mk + 64 * w), zero);
      for (int blk = 0; blk < kNB; blk++) {
        const uint16_t *q = r + blk * kB;
        for (int j = 0; j < kB; j += 4) {
          mk[q[j]] = 0x80;
          mk[q[j + 1]] = 0x80;
          mk[q[j + 2]] = 0x80;
          mk[q[j + 3]] = 0x80;
        }
        uint64_t *o = pr + (size_t)blk * W;
        for (int w = 0; w < W; w++) {
          const __m512i x = _mm512_load_si512((const void *)(mk + 64 * w));
          o[w] = (uint64_t)_mm512_movepi8_mask(x);
          _mm512_store_si512((void *)(mk + 64 * w), zero);
        }
      }
    }
#else
    static bool avx512() { return false; }
#endif
  };

  // A walk over the chain on one thread, making its own superblocks (the
  // two most recent are kept: a walk only moves forward).
  class Walker {
   public:
    explicit Walker(const PermChain &c) : c_(c), g_(c.g_) {
      miss_.assign(g_.W, 0);
      tmp_.assign(g_.W, 0);
    }
    // the permutation starting at p: its kept state, and the next start
    long long step(long long p, PermStream::State *kept) {
      const int n = g_.n, W = g_.W;
      p += n;  // discarded draws
      state_at(get(p / kL), (int)(p % kL), kept->h);
      uint64_t *miss = miss_.data();
      for (int w = 0; w < W; w++) miss[w] = ~0ull;
      if (n & 63) miss[W - 1] = (1ull << (n & 63)) - 1;
      int left = n;
      if (p % kB) {
        // the partial block up to the next block boundary: its residues only
        // leave the missing set, unless it could hold every missing one
        const Slot &s = get(p / kL);
        const int off = (int)(p % kL), bend = (off / kB + 1) * kB;
        const uint16_t *r = s.res.data();
        if (left <= bend - off) {
          int i = off;
          for (; i < bend && left > 0; i++) {
            const uint32_t v = r[i];
            const uint64_t bit = 1ull << (v & 63);
            left -= (miss[v >> 6] & bit) != 0;
            miss[v >> 6] &= ~bit;
          }
          p += i - off;
        } else {
          for (int i = off; i < bend; i++) miss[r[i] >> 6] &= ~(1ull << (r[i] & 63));
          p += bend - off;
        }
      }
      if (left > 0) {
        p = skip_blocks(p);
        p = last_first(p);
      }
      return p;
    }

   private:
    const Slot &get(long long k) {
      Slot &s = sl_[k & 1];
      if (sk_[k & 1] == k) return s;
      Vec st;
      if (sk_[(k - 1) & 1] == k - 1 && k > 0)
        st = mat_vec(jumps()[0], start_[(k - 1) & 1]);
      else
        st = c_.superblock_start(k);
      g_.fill(s, st, raw_);
      sk_[k & 1] = k;
      start_[k & 1] = st;
      return s;
    }
    static void state_at(const Slot &s, int j, uint32_t *h) {
      uint32_t b[kH + kB];
      const int blk = j / kB, o = j - blk * kB;
      for (int i = 0; i < kH; i++) b[i] = s.ckpt[(size_t)blk * kH + i];
      for (int i = kH; i < kH + o; i++) b[i] = b[i - 31] + b[i - 3];
      for (int i = 0; i < kH; i++) h[i] = b[o + i];
    }
    // From block-aligned position p: moves past every block that lacks one of
    // the residues in miss_ (removing the residues it holds) and returns the
    // start of the first block that holds all of them.
    long long skip_blocks(long long p) {
#if defined(__x86_64__)
      if (Gen::avx512()) {
        switch ((g_.W + 7) >> 3) {
          case 1: return skip_blocks_avx512<1>(p);
          case 2: return skip_blocks_avx512<2>(p);
          case 3: return skip_blocks_avx512<3>(p);
          case 4: return skip_blocks_avx512<4>(p);
          default: break;
        }
      }
#endif
      const int W = g_.W;
      uint64_t *miss = miss_.data();
      for (;;) {
        const Slot &s = get(p / kL);
        for (int b = (int)(p % kL) / kB; b < kNB; b++, p += kB) {
          const uint64_t *pr = &s.pres[(size_t)b * W];
          uint64_t lack = 0;
          for (int w = 0; w < W; w++) lack |= miss[w] & ~pr[w];
          if (!lack) return p;
          for (int w = 0; w < W; w++) miss[w] &= ~pr[w];
        }
      }
    }
    // Block-aligned p starts a block that holds every residue of miss_: the
    // position after the last of their first appearances in it.
    long long last_first(long long p) {
      const int W = g_.W;
      const uint64_t *miss = miss_.data();
      const Slot &s = get(p / kL);
      const uint16_t *r = s.res.data() + p % kL;
      int left = 0;
      for (int w = 0; w < W; w++) left += __builtin_popcountll(miss[w]);
#if defined(__x86_64__)
      if (Gen::avx512() && left <= 16) return p + last_first_avx512(r, miss, W) + 1;
#endif
      uint64_t *mm = tmp_.data();
      for (int w = 0; w < W; w++) mm[w] = miss[w];
      int i = 0;
      for (;; i++) {
        const uint32_t v = r[i];
        const uint64_t bit = 1ull << (v & 63);
        left -= (mm[v >> 6] & bit) != 0;
        mm[v >> 6] &= ~bit;
        if (left == 0) break;
      }
      return p + i + 1;
    }
#if defined(__x86_64__)
    template <int NV>
    __attribute__((target("avx512f,avx512bw"))) long long skip_blocks_avx512(long long p) {
      const int W = g_.W;
      const __mmask8 tail = (W & 7) ? (__mmask8)((1u << (W & 7)) - 1u) : (__mmask8)0xFF;
      uint64_t *miss = miss_.data();
      __m512i mv[NV];
#pragma GCC unroll 4
      for (int v = 0; v < NV; v++) mv[v] = _mm512_maskz_loadu_epi64(v == NV - 1 ? tail : 0xFF, miss + 8 * v);
      for (bool done = false; !done;) {
        const Slot &s = get(p / kL);
        for (int b = (int)(p % kL) / kB; b < kNB; b++, p += kB) {
          const uint64_t *pr = &s.pres[(size_t)b * W];
          __m512i x[NV];
          __mmask8 lack = 0;
#pragma GCC unroll 4
          for (int v = 0; v < NV; v++) {
            x[v] = _mm512_maskz_loadu_epi64(v == NV - 1 ? tail : 0xFF, pr + 8 * v);
            lack |= _mm512_test_epi64_mask(_mm512_andnot_si512(x[v], mv[v]), mv[v]);
          }
          if (!lack) {
            done = true;
            break;
          }
#pragma GCC unroll 4
          for (int v = 0; v < NV; v++) mv[v] = _mm512_andnot_si512(x[v], mv[v]);
        }
      }
#pragma GCC unroll 4
      for (int v = 0; v < NV; v++) _mm512_mask_storeu_epi64(miss + 8 * v, v == NV - 1 ? tail : 0xFF, mv[v]);
      return p;
    }
    // index of the last first appearance in r[0..kB) of the (at most 16)
    // residues in miss, each of which occurs there
    __attribute__((target("avx512f,avx512bw"))) static int last_first_avx512(const uint16_t *r, const uint64_t *miss,
                                                                              int W) {
      static_assert(kB == 128, "four 32-lane compares per block");
      const __m512i r0 = _mm512_loadu_si512((const void *)r), r1 = _mm512_loadu_si512((const void *)(r + 32));
      const __m512i r2 = _mm512_loadu_si512((const void *)(r + 64)), r3 = _mm512_loadu_si512((const void *)(r + 96));
      int last = 0;
      for (int w = 0; w < W; w++) {
        uint64_t m = miss[w];
        while (m) {
          const int v = (w << 6) + __builtin_ctzll(m);
          m &= m - 1;
          const __m512i vv = _mm512_set1_epi16((short)v);
          const uint64_t lo = (uint64_t)_mm512_cmpeq_epi16_mask(r0, vv) |
                              ((uint64_t)_mm512_cmpeq_epi16_mask(r1, vv) << 32);
          int at;
          if (lo) {
            at = __builtin_ctzll(lo);
          } else {
            const uint64_t hi = (uint64_t)_mm512_cmpeq_epi16_mask(r2, vv) |
                                ((uint64_t)_mm512_cmpeq_epi16_mask(r3, vv) << 32);
            at = 64 + __builtin_ctzll(hi);
          }
          last = at > last ? at : last;
        }
      }
      return last;
    }
#endif
    const PermChain &c_;
    const Gen &g_;
    Slot sl_[2];
    long long sk_[2] = {-1, -1};
    Vec start_[2];
    std::vector<uint32_t> raw_;
    std::vector<uint64_t> miss_, tmp_;
  };

  // A segment: chain entries [c0, c1) expected; its walk from `from` records
  // start positions and kept states until it passes `until`.
  struct Seg {
    long long c0 = 0, c1 = 0, from = 0, until = 0;
    std::vector<long long> pos;
    std::vector<PermStream::State> kept;
    long long next = 0;  // the start after the last recorded entry
    std::atomic<int> done{0};
  };

  // Segments: a short first one from the chain's start, then 256, 512 and
  // 1,024 permutations (several of each, so that many threads work near the
  // validated end from the start), then 2,048.  A walk's guess is taken when
  // its segment is claimed, from the last validated start: 4.5 standard
  // deviations of the summed lengths before the expected start; it goes on
  // past the next segment's latest likely start by kExtend permutations.
  static const int kExtend = 40;
  void plan() {
    const int n = g_.n;
    double H = 0;
    for (int k = 1; k <= n; k++) H += 1.0 / k;
    mean_ = n + n * H;                                     // draws per permutation
    sd_ = n * 3.14159265358979 / std::sqrt(6.0);           // of one permutation's length
    std::vector<long long> bounds{0};
    const long long sizes[] = {64, 256, 256, 256, 256, 512, 512, 512, 512, 1024, 1024, 1024, 1024};
    for (int i = 0; bounds.back() < count_ + 1; i++) {
      const long long len = i < (int)(sizeof(sizes) / sizeof(sizes[0])) ? sizes[i] : 2048;
      bounds.push_back(std::min(count_ + 1, bounds.back() + len));
    }
    nseg_ = (int)bounds.size() - 1;
    segs_plan_.resize(nseg_);
    for (int j = 0; j < nseg_; j++) {
      segs_plan_[j].c0 = bounds[j];
      segs_plan_[j].c1 = bounds[j + 1];
    }
  }
  struct Plan {
    long long c0, c1;
  };
  // where a walk for chain entries [c0, c1) starts and stops, judged from
  // the validated entries so far
  void guess(long long c0, long long c1, long long &from, long long &until) const {
    const long long V = validated();
    const long long v0 = V > 0 ? V - 1 : 0;          // a validated entry
    const double p0 = V > 0 ? (double)pos_[v0] : 0.0;  // its start
    auto est = [&](long long c) { return p0 + (double)(c - v0) * mean_; };
    auto lead = [&](long long c) { return 4.5 * sd_ * std::sqrt((double)std::max(0LL, c - v0)) + 2.0 * mean_; };
    from = std::max((long long)p0, (long long)(est(c0) - lead(c0)));
    until = (long long)(est(c1) + lead(c1) + kExtend * mean_);
  }
  double mean_ = 0, sd_ = 0;

  void seg_loop() {
#if defined(__linux__)
    // walks yield to the caller's threads (the GPU runtime's submissions,
    // the chunk bookkeeping): they only need the cores left over
    if (nice_ > 0) setpriority(PRIO_PROCESS, (id_t)syscall(SYS_gettid), nice_);
#endif
    Walker w(*this);
    for (;;) {
      const int j = next_seg_.fetch_add(1);
      if (j >= nseg_ || quit_.load(std::memory_order_relaxed)) return;
      Seg &s = segs_[j];
      const Plan &q = segs_plan_[j];
      s.c0 = q.c0;
      s.c1 = q.c1;
      if (j == 0) {
        s.from = 0;
        s.until = 0;
      } else {
        guess(q.c0, q.c1, s.from, s.until);
      }
      long long p = s.from;
      // segment 0 starts at the chain's start: it needs only its own entries
      const long long cap = j == 0 ? q.c1 : -1;
      while (!quit_.load(std::memory_order_relaxed)) {
        PermStream::State k;
        const long long nx = w.step(p, &k);
        s.pos.push_back(p);
        s.kept.push_back(k);
        p = nx;
        if (cap >= 0 ? (long long)s.pos.size() >= cap : p > s.until) break;
      }
      s.next = p;
      {
        std::lock_guard<std::mutex> g(mu_);
        s.done.store(1, std::memory_order_release);
      }
      cv_.notify_all();
    }
  }

  // appends entries to the validated chain (the merger thread only)
  void publish(long long v) {
    {
      std::lock_guard<std::mutex> g(mu_);
      valid_.store(v, std::memory_order_release);
    }
    vcv_.notify_all();
  }

  void merge_loop() {
    Walker w(*this);
    long long V = 0;          // validated entries
    long long tail_next = 0;  // the start of entry V (after entry V - 1's permutation)
    auto walk_one = [&]() {
      PermStream::State k;
      const long long nx = w.step(tail_next, &k);
      pos_[V] = tail_next;
      kept_[V] = k;
      V++;
      fallback_++;
      tail_next = nx;
    };
    for (int j = 0; j < nseg_ && V <= count_; j++) {
      Seg &s = segs_[j];
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [&]() { return s.done.load(std::memory_order_acquire) || quit_.load(std::memory_order_relaxed); });
      }
      if (quit_.load(std::memory_order_relaxed)) return;
      const std::vector<long long> &T = s.pos;
      const long long nT = (long long)T.size();
      long long take = -1;  // index in T of validated entry V
      if (j == 0) {
        take = 0;  // the walk from the chain's start
      } else if (nT > 0) {
        // the first validated start the walk also passed: from there on they
        // are one chain, so entry V lies as far past it in T as in the chain
        long long a = std::lower_bound(pos_.begin(), pos_.begin() + V, T.front()) - pos_.begin();
        for (; a < V && take < 0; a++) {
          auto it = std::lower_bound(T.begin(), T.end(), pos_[a]);
          if (it != T.end() && *it == pos_[a]) take = (it - T.begin()) + (V - a);
        }
        if (take < 0) {
          // walk on from the validated end until a start the walk passed
          for (;;) {
            auto it = std::lower_bound(T.begin(), T.end(), tail_next);
            if (it != T.end() && *it == tail_next) {
              take = it - T.begin();
              break;
            }
            if (tail_next > T.back() || V > count_ || quit_.load(std::memory_order_relaxed)) break;
            walk_one();
          }
        }
      }
      if (take >= 0 && take <= nT) {
        long long b = take;
        for (; b < nT && V <= count_; b++, V++) {
          pos_[V] = T[b];
          kept_[V] = s.kept[b];
        }
        // entry V - 1 = T[b - 1] (also when nothing was appended and take is
        // one past T's end): the start after it
        if (b > take || take == nT) tail_next = b < nT ? T[b] : s.next;
      }
      publish(V);
      std::vector<long long>().swap(s.pos);
      std::vector<PermStream::State>().swap(s.kept);
    }
    // entries no segment reached: walked here
    while (V <= count_ && !quit_.load(std::memory_order_relaxed)) {
      walk_one();
      if ((V & 63) == 0) publish(V);
    }
    publish(V);
  }

  Gen g_;
  long long count_;
  int nice_ = 0;
  Vec base_;
  std::vector<long long> pos_;
  std::vector<PermStream::State> kept_;
  int nseg_ = 0;
  std::vector<Plan> segs_plan_;
  std::unique_ptr<Seg[]> segs_;
  std::atomic<int> next_seg_{0};
  std::atomic<long long> valid_{0};
  std::atomic<long long> fallback_{0};
  std::atomic<long long> wait_us_{0};
  std::atomic<bool> quit_{false};
  std::mutex mu_;
  std::condition_variable cv_, vcv_;
  std::vector<std::thread> pool_;
  std::thread merger_;
};
