// perm_index.h - the permutation stream of perm_stream.h, consumed with the
// help of per-block presence bitsets computed ahead on worker threads.
//
// One permutation's consumption (main.cu:541-558) is: n discarded draws, the
// draws up to the n1-th distinct residue (class 1), then the draws up to the
// first appearance of every other residue.  Scanned draw by draw that is
// about n * H(n) steps (4,200 for n = 600), and the stream is one sequential
// chain.  Two facts split that work:
//   - where a permutation ends depends only on where it starts: the end is
//     the first position at which every residue has appeared since the
//     discarded draws (the class split does not move it);
//   - the class split depends only on the stream state at the first kept
//     draw (34 values), and needs only the draws up to the n1-th distinct
//     residue (about n ln(n / (n - n1)): 416 for 300 of 600).
// So one thread (the chain) only finds each permutation's end and records the
// state at its first kept draw; the split is recomputed from that state by
// any number of other threads (`class2_from_state`), in any order.
//
// For the chain, the stream is cut into superblocks whose starting states are
// jumped to directly (the recurrence is linear over Z/2^32, so L steps are one
// 34x34 matrix power), and worker threads produce, for each superblock, the
// raw values, their residues mod n, and for every block of kB draws the set of
// residues present in it.  The chain then takes whole blocks while a block
// lacks one of the residues still missing (removing the ones it holds: a few
// vector operations per block), and finds the end inside the first block that
// holds them all by comparing its residues with each one still missing.  The
// results (class-2 bits, stream position) are exactly those of
// PermStream::class2_bits.
#pragma once
#include <stdint.h>

#include <atomic>
#include <chrono>
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
  static const int kC = 256;       // draws generated at a time by class2_from_state

  // Per-thread working memory of class2_from_state.
  struct Scratch {
    std::vector<uint32_t> raw;  // kH values of history, then kC values
    std::vector<uint16_t> res;  // residues of the kC values
    std::vector<uint64_t> c1;   // class-1 residues so far
  };

  // Starts at the stream state `st` (the position of the next draw); every
  // permutation has n samples of which n1 are class 1.
  PermIndex(const PermStream::State &st, int n, int n1, int workers) : n_(n), n1_(n1) {
    W_ = (n + 63) / 64;
    miss_.assign(W_, 0);
    miss_tmp_.assign(W_, 0);
    K_ = workers < 1 ? 1 : workers;
    R_ = 4 * K_;
    sh_ = 0;
    while ((2 << sh_) < n) sh_++;  // 2^sh < n <= 2^(sh+1)
    m_ = n > 1 ? ((uint64_t)1 << (32 + sh_)) / (uint32_t)n + 1 : 0;
    slots_.reset(new Slot[R_]);
    for (int i = 0; i < R_; i++) {
      slots_[i].ckpt.resize((size_t)kNB * kH);
      slots_[i].res.resize(kL);
      slots_[i].pres.resize((size_t)kNB * W_);
    }
    // workers claim superblocks in order from a shared counter; the starting
    // state of superblock k + 1 is published by the claimer of k
    states_.reset(new StateSlot[R_]);
    for (int i = 0; i < kH; i++) states_[0].v.v[i] = st.h[i];
    states_[0].k.store(0, std::memory_order_release);
    for (int w = 0; w < K_; w++) pool_.emplace_back([this]() { work(); });
  }
  ~PermIndex() {
    quit_.store(true);
    for (auto &t : pool_) t.join();
  }

  // The chain: one permutation's draws.  `st` (when given) receives the
  // stream state at its first kept draw (after the n discarded ones), from
  // which class2_from_state recomputes its class split; the position moves
  // past the permutation's draws.
  void skip_perm(PermStream::State *st) {
    const int n = n_, W = W_;
    long long p = g_ + n;  // discarded draws
    if (st) state_at(get(p / kL), (int)(p % kL), st->h);
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
      // whole blocks while a block lacks a missing residue; then the first
      // appearances of the residues still missing inside the block that
      // holds them all
      p = skip_blocks(p);
      p = last_first(p);
    }
    g_ = p;
  }

  // The class-2 bits (c2, (n + 31) / 32 words) of the permutation whose first
  // kept draw follows the stream state st: class 1 is the first n1 distinct
  // residues rand() % n (main.cu:541-570).  Reads nothing shared: any thread,
  // any order.
  void class2_from_state(const PermStream::State &st, uint32_t *c2, Scratch &sc) const {
    const int n = n_, n1 = n1_, W = W_;
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
      residues(b + kH, r, kC);
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

  // time the chain has waited for superblocks
  double wait_ms() const { return wait_ms_; }

  // The stream state at the current position (the 34 values before it).
  PermStream::State save() {
    PermStream::State st;
    state_at(get(g_ / kL), (int)(g_ % kL), st.h);
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
    std::vector<uint32_t> ckpt;    // per block: the kH values before it
    std::vector<uint16_t> res;     // (value >> 1) mod n
    std::vector<uint64_t> pres;    // per block: residues present
  };

  // h = the kH values before position j of the slot's superblock: from the
  // checkpoint of j's block, up to 127 steps of the recurrence
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
    if (avx512()) {
      switch ((W_ + 7) >> 3) {
        case 1: return skip_blocks_avx512<1>(p);
        case 2: return skip_blocks_avx512<2>(p);
        case 3: return skip_blocks_avx512<3>(p);
        case 4: return skip_blocks_avx512<4>(p);
        default: break;
      }
    }
#endif
    const int W = W_;
    uint64_t *miss = miss_.data();
    for (;;) {
      const Slot &s = get(p / kL);
      for (int b = (int)(p % kL) / kB; b < kNB; b++, p += kB) {
        ahead(s, b);
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
    const int W = W_;
    const uint64_t *miss = miss_.data();
    const Slot &s = get(p / kL);
    const uint16_t *r = s.res.data() + p % kL;
    int left = 0;
    for (int w = 0; w < W; w++) left += __builtin_popcountll(miss[w]);
#if defined(__x86_64__)
    if (avx512() && left <= 16) return p + last_first_avx512(r, miss, W) + 1;
#endif
    uint64_t *mm = miss_tmp_.data();
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
    const int W = W_;
    const __mmask8 tail = (W & 7) ? (__mmask8)((1u << (W & 7)) - 1u) : (__mmask8)0xFF;
    uint64_t *miss = miss_.data();
    __m512i mv[NV];
#pragma GCC unroll 4
    for (int v = 0; v < NV; v++) mv[v] = _mm512_maskz_loadu_epi64(v == NV - 1 ? tail : 0xFF, miss + 8 * v);
    for (bool done = false; !done;) {
      const Slot &s = get(p / kL);
      for (int b = (int)(p % kL) / kB; b < kNB; b++, p += kB) {
        ahead(s, b);
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

  // The chain reads what other cores wrote: ask for the presence words some
  // blocks ahead before they are needed.
  inline void ahead(const Slot &s, int b) const {
    if (b + 16 < kNB) {
      const char *q = (const char *)&s.pres[(size_t)(b + 16) * W_];
      __builtin_prefetch(q, 0, 3);
      __builtin_prefetch(q + 8 * W_ - 1, 0, 3);
    }
  }

  // superblock k, once a worker has filled it; the superblocks before k are
  // then no longer read (the chain only moves forward, and the state saved at
  // the end of a permutation lies in the superblock of its position)
  const Slot &get(long long k) {
    if (k > release_.load(std::memory_order_relaxed)) release_.store(k, std::memory_order_release);
    Slot &s = slots_[k % R_];
    if (s.k.load(std::memory_order_acquire) != k) {
      const auto w0 = std::chrono::steady_clock::now();
      int spins = 0;
      while (s.k.load(std::memory_order_acquire) != k) {
        if (++spins > 64) std::this_thread::yield();
      }
      wait_ms_ += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    }
    return s;
  }

  // A worker claims the next superblock once its slot is free (the chain has
  // moved past the superblock R_ before it), takes its starting state,
  // publishes the next one's (one 34x34 step of M^kL) and fills the slot.
  // Claims go in order, so however the threads are scheduled the lowest
  // superblocks are filled first.
  void work() {
    const Mat &P = jump_superblock();
    std::vector<uint32_t> raw;  // the scalar path's values
    for (;;) {
      long long k = next_.load(std::memory_order_relaxed);
      int spins = 0;
      for (;;) {
        if (quit_.load(std::memory_order_relaxed)) return;
        if (k - R_ < release_.load(std::memory_order_acquire) &&
            next_.compare_exchange_weak(k, k + 1, std::memory_order_acq_rel))
          break;
        if (++spins > 64) std::this_thread::yield();
        k = next_.load(std::memory_order_relaxed);
      }
      StateSlot &cur = states_[k % R_];
      spins = 0;
      while (cur.k.load(std::memory_order_acquire) != k) {
        if (quit_.load(std::memory_order_relaxed)) return;
        if (++spins > 64) std::this_thread::yield();
      }
      const Vec st = cur.v;
      StateSlot &nx = states_[(k + 1) % R_];
      nx.v = mat_vec(P, st);
      nx.k.store(k + 1, std::memory_order_release);
      Slot &s = slots_[k % R_];
      fill(s, st, raw);
      s.k.store(k, std::memory_order_release);
    }
  }

  // A superblock from its starting state: the residues of its values, the
  // checkpoint of each block, the presence words of each block.
  void fill(Slot &s, const Vec &st, std::vector<uint32_t> &raw) {
#if defined(__x86_64__)
    if (avx512() && n_ > 1) {
      values_avx512(st, s);
    } else
#endif
    {
      raw.resize(kH + kL);
      uint32_t *b = raw.data();
      for (int i = 0; i < kH; i++) b[i] = st.v[i];
      recurrence(b);
      residues(b + kH, s.res.data(), kL);
      for (int blk = 0; blk < kNB; blk++)
        for (int i = 0; i < kH; i++) s.ckpt[(size_t)blk * kH + i] = b[blk * kB + i];
    }
    uint64_t *pr = s.pres.data();
    const uint16_t *r = s.res.data();
#if defined(__x86_64__)
    if (avx512() && W_ <= 32) {
      presence_avx512(r, pr);
      return;
    }
#endif
    std::memset(pr, 0, sizeof(uint64_t) * (size_t)kNB * W_);
    for (int blk = 0; blk < kNB; blk++) {
      uint64_t *q = pr + (size_t)blk * W_;
      for (int j = blk * kB; j < (blk + 1) * kB; j++) q[r[j] >> 6] |= 1ull << (r[j] & 63);
    }
  }
  // b[kH .. kH + kL) from the kH values before them: three independent
  // chains (r[i] depends on r[i-3])
  static void recurrence(uint32_t *b) {
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
  }
#if defined(__x86_64__)
  // The same values 16 at a time, turned into residues as they are made, with
  // the checkpoints taken from the vectors in registers (no array of values).
  // Substituting r[i-3] = r[i-6] + r[i-34] five times gives r[i] = r[i-18] +
  // r[i-31] + r[i-34] + r[i-37] + r[i-40] + r[i-43] + r[i-46]: every term lies
  // at least 18 back, so a vector of 16 consecutive values is a sum of windows
  // of the three vectors before it.
  __attribute__((target("avx512f,avx512bw"))) void values_avx512(const Vec &st, Slot &s) const {
    static_assert(kH == 34 && kL % kB == 0 && kB % 16 == 0, "layout");
    // positions -34 .. 13 (the history, then 14 values one at a time), so
    // that the vectors start at position 14 and a block's start, a multiple
    // of 128, is element 2 of a vector
    alignas(64) uint32_t b[48];
    for (int i = 0; i < kH; i++) b[i] = st.v[i];
    for (int i = kH; i < 48; i++) b[i] = b[i - 31] + b[i - 3];
    for (int i = 0; i < kH; i++) s.ckpt[i] = b[i];
    uint16_t *res = s.res.data();
    for (int i = 0; i < 14; i++) res[i] = mod_n(b[kH + i] >> 1);
    __m512i p3 = _mm512_load_si512((const void *)b), p2 = _mm512_load_si512((const void *)(b + 16));
    __m512i p1 = _mm512_load_si512((const void *)(b + 32));
    const __m512i vm = _mm512_set1_epi64((long long)m_);
    const __m512i vn = _mm512_set1_epi32(n_);
    const __m128i shv = _mm_cvtsi32_si128(32 + sh_);
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
      // residues (value >> 1) mod n, as residues_avx512
      const __m512i x = _mm512_srli_epi32(v, 1);
      const __m512i qe = _mm512_srl_epi64(_mm512_mul_epu32(x, vm), shv);
      const __m512i qo = _mm512_srl_epi64(_mm512_mul_epu32(_mm512_srli_epi64(x, 32), vm), shv);
      const __m512i qq = _mm512_or_si512(_mm512_and_si512(qe, lo32), _mm512_slli_epi64(qo, 32));
      _mm256_storeu_si256((__m256i *)(res + q), _mm512_cvtepi32_epi16(_mm512_sub_epi32(x, _mm512_mullo_epi32(qq, vn))));
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
  // cleared (no read-modify-write of the words per draw).
  __attribute__((target("avx512f,avx512bw"))) void presence_avx512(const uint16_t *r, uint64_t *pr) const {
    alignas(64) uint8_t m[32 * 64];
    const int W = W_;
    const __m512i zero = _mm512_setzero_si512();
    for (int w = 0; w < W; w++) _mm512_store_si512((void *)(m + 64 * w), zero);
    for (int blk = 0; blk < kNB; blk++) {
      const uint16_t *q = r + blk * kB;
      for (int j = 0; j < kB; j += 4) {
        m[q[j]] = 0x80;
        m[q[j + 1]] = 0x80;
        m[q[j + 2]] = 0x80;
        m[q[j + 3]] = 0x80;
      }
      uint64_t *o = pr + (size_t)blk * W;
      for (int w = 0; w < W; w++) {
        const __m512i x = _mm512_load_si512((const void *)(m + 64 * w));
        o[w] = (uint64_t)_mm512_movepi8_mask(x);
        _mm512_store_si512((void *)(m + 64 * w), zero);
      }
    }
  }
#endif
  inline uint16_t mod_n(uint32_t x) const {
    if (n_ <= 1) return 0;
    const uint32_t q = (uint32_t)(((uint64_t)x * m_) >> (32 + sh_));
    return (uint16_t)(x - q * (uint32_t)n_);
  }
  // residues of `count` values (a multiple of 16)
  void residues(const uint32_t *raw, uint16_t *res, int count) const {
#if defined(__x86_64__)
    if (n_ > 1 && avx512()) {
      residues_avx512(raw, res, count);
      return;
    }
#endif
    for (int i = 0; i < count; i++) res[i] = mod_n(raw[i] >> 1);
  }
#if defined(__x86_64__)
  static bool avx512() {
    static const bool ok = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw");
    return ok;
  }
  // the quotient (x m) >> (32 + s) in 64-bit lanes, as in perm_stream.h
  __attribute__((target("avx512f,avx512bw"))) void residues_avx512(const uint32_t *raw, uint16_t *res,
                                                                  int count) const {
    const __m512i vm = _mm512_set1_epi64((long long)m_);
    const __m512i vn = _mm512_set1_epi32(n_);
    const __m128i shv = _mm_cvtsi32_si128(32 + sh_);
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
#endif

  int n_, n1_, W_, K_, R_;
  std::vector<uint64_t> miss_, miss_tmp_;  // the residues still missing, scratch (W_ words each)
  int sh_ = 0;
  uint64_t m_ = 0;
  long long g_ = 0;  // draws consumed since the starting state
  double wait_ms_ = 0;
  struct StateSlot {
    std::atomic<long long> k{-1};  // superblock whose starting state v holds, or -1
    Vec v;
  };
  std::unique_ptr<Slot[]> slots_;
  std::unique_ptr<StateSlot[]> states_;
  std::atomic<long long> next_{0};     // the next superblock to claim
  std::atomic<long long> release_{0};  // superblocks below this one are no longer read
  std::atomic<bool> quit_{false};
  std::vector<std::thread> pool_;
};
