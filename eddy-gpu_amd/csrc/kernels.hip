// kernels.hip - gfx950 kernels of the EDDY permutation loop.
//
// Work unit: an "item" = (gene set, permutation).  For every item the engine
// produces, exactly as the reference's per-permutation launches do
// (main.cu:656-1428):
//   k_count_test  full-class 3x3 tables of every gene pair as an int8 one-hot
//                 GEMM on MFMA (v_mfma_i32_16x16x64_i8, int32 accumulate),
//                 then the chi-square/AS 147 edge tests (kernel1.cu:10-477)
//   k_loo         leave-one-out tests up to each row's cutoff (cutoff mode)
//   k_rows        capped parent lists of the two full networks, the rows of
//                 each leave-one-out network that differ from them, and a
//                 64-bit hash per network (kernel2.cu:114, kernel3.cu:41-165)
//   k_score       first-occurrence dedup with exact confirmation
//                 (kernel4.cu, main.cu:1060-1121), BDeu family histograms
//                 (kernel5.cu), scaled softmax likelihoods and JS divergence
//                 (main.cu:1294-1428)
//   k_pvalue      beta fit + ibetac per set (main.cu:1462-1502)
//
// Leave-one-out identity used throughout: the tally of network s (class g,
// sample k left out) is the class-g full tally minus one count in the cell
// (x_row[k], x_col[k]) (kernel1.cu:86-119 skips exactly that sample), so each
// (pair, class) has at most 10 distinct tables and 10 distinct decisions.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine_types.h"
#include "edgetest.h"
#include "fpmath.h"

typedef int v4i __attribute__((ext_vector_type(4)));

// chi-square edge decision of one 3x3 tally: edgetest.h (kernel1.cu:10-203,
// 420-431 with an exact shortcut of the AS 147 series)
__device__ __forceinline__ int d_edge_test(const int *t, int prior, const KArgs &a, const SetMeta &sm,
                                           double cpairs) {
  return eddy_edge_decision_fast(t, prior, &a.ep, &sm.eb, cpairs);
}

// Leave-one-out decisions of one (pair, class): bit cell set when the tally
// minus one count in `cell` is an edge; empty cells copy the full decision.
__device__ __forceinline__ uint32_t d_loo_mask(const int *t, int full, int prior, const KArgs &a,
                                               const SetMeta &sm, double cpairs, int *ntests) {
  return eddy_loo_mask_fast(t, full, prior, &a.ep, &sm.eb, cpairs, ntests);
}

__device__ __forceinline__ long long pair_index(int t, int col) { return (long long)t * (t - 1) / 2 + col; }

// A class's 3x3 tally as packed u16 counts: cells 0..7 in two 8-byte words,
// cell 8 in a third only when the dataset stores it (a.tabw ==
// EDDY_TAB_FULL); otherwise it is the class size minus the other eight.
__device__ __forceinline__ void store_tally(uint16_t *dst, const int *t, int tabw) {
  const uint64_t w0 = (uint64_t)(uint16_t)t[0] | ((uint64_t)(uint16_t)t[1] << 16) |
                      ((uint64_t)(uint16_t)t[2] << 32) | ((uint64_t)(uint16_t)t[3] << 48);
  const uint64_t w1 = (uint64_t)(uint16_t)t[4] | ((uint64_t)(uint16_t)t[5] << 16) |
                      ((uint64_t)(uint16_t)t[6] << 32) | ((uint64_t)(uint16_t)t[7] << 48);
  if (tabw == EDDY_TAB_DENSE) {
    *(ulonglong2 *)dst = make_ulonglong2(w0, w1);  // 16-byte aligned: one store
  } else {
    uint64_t *d = (uint64_t *)dst;
    d[0] = w0;
    d[1] = w1;
    d[2] = (uint64_t)(uint16_t)t[8];
  }
}
// cell 8 of a tally of n samples from its cells 0..7 (packed in q0, q1): every
// partial sum is at most n <= 20000, so the packed additions never carry
// across a 16-bit field
__device__ __forceinline__ uint64_t tally_cell8(uint64_t q0, uint64_t q1, int n) {
  uint64_t s = q0 + q1;
  s += s >> 32;
  s += s >> 16;
  return (uint64_t)(uint32_t)(n - (int)(s & 0xffffu));
}
// the tally words (q0, q1, q2) of the table at q: q2 loaded, or derived from
// the n samples the table counts
__device__ __forceinline__ void load_tally_words(const uint16_t *q, int tabw, int n, uint64_t &q0, uint64_t &q1,
                                                 uint64_t &q2) {
  if (tabw == EDDY_TAB_DENSE) {
    const ulonglong2 v = *(const ulonglong2 *)q;
    q0 = v.x;
    q1 = v.y;
    q2 = tally_cell8(q0, q1, n);
  } else {
    const uint64_t *w = (const uint64_t *)q;
    q0 = w[0];
    q1 = w[1];
    q2 = w[2];
  }
}
// class 1's tally of pair p of the item (cutoff schedule)
__device__ __forceinline__ uint16_t *tally_at(const KArgs &a, const ItemMeta &im, long long p) {
  return a.tab + (size_t)(im.infooff + p) * a.tabw;
}
// the pooled table of pair p of set sm
__device__ __forceinline__ const uint16_t *totals_at(const KArgs &a, const SetMeta &sm, long long p) {
  return a.totals + (size_t)(sm.pairbase + p) * a.tabw;
}

// the packed tally words of pair p, class g: class 1's as stored; class 2's
// as the pair's pooled table minus class 1's, word by word (every pooled
// count is at least class 1's, so no borrow crosses a 16-bit field)
__device__ __forceinline__ void class_tally_words(const KArgs &a, const ItemMeta &im, const SetMeta &sm, long long p,
                                                  int g, uint64_t &q0, uint64_t &q1, uint64_t &q2) {
  load_tally_words(tally_at(a, im, p), a.tabw, a.n1, q0, q1, q2);
  if (g) {
    uint64_t t0, t1, t2;
    load_tally_words(totals_at(a, sm, p), a.tabw, a.N, t0, t1, t2);
    q0 = t0 - q0;
    q1 = t1 - q1;
    q2 = t2 - q2;
  }
}

__device__ __forceinline__ void unpack_tally(uint64_t q0, uint64_t q1, uint64_t q2, int *t) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
    t[i] = (int)((q0 >> (16 * i)) & 0xffff);
    t[4 + i] = (int)((q1 >> (16 * i)) & 0xffff);
  }
  t[8] = (int)(q2 & 0xffff);
}

__device__ __forceinline__ void load_class_tally(const KArgs &a, const ItemMeta &im, const SetMeta &sm, long long p,
                                                 int g, int *t) {
  uint64_t q0, q1, q2;
  class_tally_words(a, im, sm, p, g, q0, q1, q2);
  unpack_tally(q0, q1, q2, t);
}

__device__ __forceinline__ int prior_bit(const KArgs &a, const SetMeta &sm, long long p) {
  return (a.prior[sm.poff + (p >> 5)] >> (p & 31)) & 1;
}

// adds v to statistics counter k (this workgroup's copy)
__device__ __forceinline__ void cnt_add(const KArgs &a, int k, unsigned long long v) {
  atomicAdd(&a.counters[(blockIdx.x & (EDDY_CNT_COPIES - 1)) * EDDY_NCNT + k], v);
}

// per-phase workgroup time (constant-rate wall clock), summed over items into
// counter 16 + phase when the engine collects statistics
__device__ __forceinline__ void phase_mark(const KArgs &a, int phase, unsigned long long &t0) {
  if (!a.counters || threadIdx.x != 0) return;
  unsigned long long t = wall_clock64();
  if (phase >= 0) cnt_add(a, 16 + phase, t - t0);
  t0 = t;
}

// LDS ordering among the lanes of one wave (a wave executes its LDS accesses
// in order; this keeps the compiler from moving them across the point)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Learned-proposal features of a pair from its pooled table (both classes of
// permutation 0 = all samples; it does not change under permutation):
// Pearson chi2/n, log1p chi2, mutual information, the two marginal entropies,
// fraction of empty cells, smallest expected count/(n/9), n/1000, dof/4,
// prior flag, log(c)/12, Cramer's V
__device__ void pair_features(const int *t1, const int *t2, int prior, int c, float *out) {
  float t[9], n = 0.f, r[3] = {0, 0, 0}, k[3] = {0, 0, 0};
  for (int i = 0; i < 9; i++) {
    t[i] = (float)(t1[i] + t2[i]);
    n += t[i];
    r[i / 3] += t[i];
    k[i % 3] += t[i];
  }
  float chi = 0.f, mi = 0.f, hr = 0.f, hk = 0.f, zeros = 0.f, emin = 1e30f;
  int nr = 0, nk = 0;
  for (int i = 0; i < 3; i++) {
    if (r[i] > 0) {
      nr++;
      hr -= (r[i] / n) * __logf(r[i] / n);
    }
    if (k[i] > 0) {
      nk++;
      hk -= (k[i] / n) * __logf(k[i] / n);
    }
  }
  for (int i = 0; i < 9; i++) {
    float e = r[i / 3] * k[i % 3] / n;
    zeros += t[i] == 0.f;
    if (e > 0.f) {
      chi += (t[i] - e) * (t[i] - e) / e;
      emin = fminf(emin, e);
      if (t[i] > 0.f) mi += (t[i] / n) * __logf(t[i] / e);
    }
  }
  if (n <= 0.f) n = 1.f;
  int mn = min(nr, nk);
  out[0] = chi / n;
  out[1] = log1pf(chi);
  out[2] = mi;
  out[3] = hr;
  out[4] = hk;
  out[5] = zeros / 9.f;
  out[6] = emin < 1e29f ? emin / (n / 9.f) : 0.f;
  out[7] = n / 1000.f;
  out[8] = (float)((nr - 1) * (nk - 1)) / 4.f;
  out[9] = (float)prior;
  out[10] = __logf((float)max(c, 1)) / 12.f;
  out[11] = mn > 1 ? sqrtf(chi / (n * (float)(mn - 1))) : 0.f;
}

// --------------------------------------------------------------------------
// K1: tables on MFMA + tests.  One workgroup = one item x one 32x32 block of
// gene pairs (row block bi >= column block bj).  One-hot rows are gene*3+state
// (96 per block); class masking is folded into the A operand, so one pass
// over all N samples accumulates both classes' tables.
// --------------------------------------------------------------------------
#define K1_TB 32
#define K1_KCH 128
#define K1_ROWS (2 * K1_TB)
// LDS strides padded off multiples of the bank width: staged operand rows
// K1_KCH + 16 bytes apart (16 rows read by one MFMA operand load land on
// distinct banks), tally rows K1_TS u16 apart (the 4 row groups of an
// accumulator store land on distinct banks)
#define K1_RS (K1_KCH + 16)
#define K1_TS 112
#define K1_OPS (3 * 96 * K1_RS)        // staged operands: A class 1, A class 2, B
#define K1_TABS (2 * 96 * K1_TS * 2)   // the two classes' tables (u16)

// Tile of workgroup `bid` of `n`.  Workgroups are dispatched to the 8 XCDs in
// turn (bid mod 8); with `grouped`, XCD x takes the x-th contiguous run of the
// tile list instead of every 8th tile, so the tiles of one item (and of the
// items of one set, which share gene rows) meet in one XCD's L2.  A bijection
// of 0..n-1: XCD x has q + (x < r) workgroups, q = n / 8, r = n mod 8.
__device__ __forceinline__ unsigned tile_of(unsigned bid, unsigned n, int grouped) {
  if (!grouped) return bid;
  const unsigned x = bid & 7u, k = bid >> 3, q = n >> 3, r = n & 7u;
  return x * q + (x < r ? x : r) + k;
}

// one-hot bytes of four samples' states (0, 1, 2; 3 = missing matches none)
// from the two low bits of each byte
__device__ __forceinline__ void one_hot3(uint32_t x, uint32_t &o0, uint32_t &o1, uint32_t &o2) {
  const uint32_t b0 = x & 0x01010101u, b1 = (x >> 1) & 0x01010101u;
  const uint32_t n0 = b0 ^ 0x01010101u, n1 = b1 ^ 0x01010101u;
  o0 = n0 & n1;
  o1 = b0 & n1;
  o2 = n0 & b1;
}

// MODE: K1_BOTH both classes on MFMA over all samples (each masked to its
// class) and every leave-one-out test here too (the all-tests schedule);
// K1_CLASS1 class 1 on MFMA over the item's class-1 samples alone (the rows
// compacted by k_gather_class1, so the contraction is n_A long, not N), class
// 2's table = the pair's pooled table minus class 1's (the cutoff schedule);
// K1_POOLED the pooled table of every pair over all samples, written to
// a.totals (once per set and run: it does not depend on the labels)
#define K1_BOTH 0
#define K1_CLASS1 1
#define K1_POOLED 2
template <int MODE>
__device__ __forceinline__ void count_test_body(const KArgs &a) {
  const uint2 w = ((const uint2 *)a.tiles)[tile_of(blockIdx.x, gridDim.x, a.tile_order)];
  const int item = (int)w.x;
  const int bi = (int)(w.y >> 16), bj = (int)(w.y & 0xffff);
  const ItemMeta im = a.items[item];
  const SetMeta &sm = a.sets[im.set];  // read in place: the bounds are indexed per tally
  const int G = sm.G, N = a.N;
  // contraction length and row stride: class 1's compacted samples, or all
  const int NP = MODE == K1_CLASS1 ? a.kp : (N + 63) & ~63;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  // the staged MFMA operands of one sample chunk, then (after the last chunk)
  // the tables: opA1[96][K1_RS] (row block: class 1, or every sample), opA2
  // (class 2, K1_BOTH only), opB (column block), row = gene * 3 + state
  constexpr int NOPS = MODE == K1_BOTH ? 3 : 2, NTAB = MODE == K1_BOTH ? 2 : 1;
  constexpr int OPS_B = NOPS * 96 * K1_RS, TAB_B = NTAB * 96 * K1_TS * 2;
  __shared__ __align__(16) uint8_t rows[OPS_B > TAB_B ? OPS_B : TAB_B];
  __shared__ int tcount;
  // full tests the screen leaves open (class-1 schedule): (pair << 1 | class),
  // decided after the tile's pairs by one lane each, so the double-precision
  // path does not hold a whole wave per undecided lane
  __shared__ uint16_t undec[2 * K1_TB * K1_TB];
  __shared__ int nundec;
  uint8_t *opA1 = rows, *opA2 = rows + 96 * K1_RS, *opB = rows + (NOPS - 1) * 96 * K1_RS;

  // class-1 membership of samples k..k+15 (k a multiple of 16) as one byte
  // per sample, from the label bits (bit = class 2); samples past N are in
  // neither class (their states are 3, one-hot zero, anyway)
  const uint32_t *lab = a.labels + im.laboff;
  auto class1_bytes = [&](int k) {
    uint32_t b = k < N ? ~(lab[k >> 5] >> (k & 31)) & 0xffffu : 0u;
    if (k + 16 > N && k < N) b &= (1u << (N - k)) - 1u;
    // bit i of a nibble to byte i: the four shifted copies do not overlap
    const uint32_t spread = 0x00204081u;
    return make_uint4((b & 15u) * spread & 0x01010101u, ((b >> 4) & 15u) * spread & 0x01010101u,
                      ((b >> 8) & 15u) * spread & 0x01010101u, ((b >> 12) & 15u) * spread & 0x01010101u);
  };
  if (tid == 0) tcount = nundec = 0;
  unsigned long long tph = 0;
  phase_mark(a, -1, tph);
  // what the tests read besides the tallies, loaded once while the operands
  // stage: the tile's prior bits (bit lj of row li = pair (gi, bj * 32 + lj)),
  // the rows' learned windows and the set's certainty intervals
  __shared__ uint32_t sh_pb[K1_TB];
  __shared__ int sh_win[K1_TB];
  __shared__ eddy_edge_bounds sh_eb;
  __shared__ eddy_screen_thr sh_thr[2];  // the screen's thresholds for class 1's and class 2's sizes
  if (tid < K1_TB) {
    const int gi = bi * K1_TB + tid;
    uint32_t pb = 0u;
    int wv = 0x7fffffff;
    if (gi < G) {
      const int nval = min(K1_TB, gi - bj * K1_TB);  // columns gj < gi of the tile
      if (nval > 0) {
        // bits base .. base + nval - 1, all of them pairs of the set
        const long long base = pair_index(gi, bj * K1_TB);
        const uint32_t *pw = a.prior + sm.poff + (base >> 5);
        const int sh = (int)(base & 31);
        uint64_t v = pw[0];
        if (sh + nval > 32) v |= (uint64_t)pw[1] << 32;
        pb = (uint32_t)(v >> sh);
      }
      if (a.win) wv = a.win[sm.genebase + gi];
    }
    sh_pb[tid] = pb;
    sh_win[tid] = wv;
  } else if (tid >= 64 && tid < 64 + 12) {
    ((double *)&sh_eb)[tid - 64] = ((const double *)&sm.eb)[tid - 64];
  } else if (MODE != K1_POOLED && tid >= 128 && tid < 128 + (int)(2 * sizeof(eddy_screen_thr) / 4)) {
    static_assert(sizeof(eddy_screen_thr) % 4 == 0, "word copy");
    const int w = tid - 128;
    uint32_t v = 0xffffffffu;  // n = -1 in the last word: no thresholds, decide nothing
    if (a.thr) v = ((const uint32_t *)(a.thr + 2 * im.set))[w];
    else if (w % (sizeof(eddy_screen_thr) / 4) != sizeof(eddy_screen_thr) / 4 - 1) v = 0u;
    ((uint32_t *)sh_thr)[w] = v;
  }

  v4i acc1[3][3], acc2[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      acc1[i][j] = (v4i){0, 0, 0, 0};
      acc2[i][j] = (v4i){0, 0, 0, 0};
    }

  const uint8_t *xs = MODE == K1_CLASS1 ? a.xc + (size_t)im.goff * a.kp : a.xs + sm.xoff;
  const int rt0 = (wave >> 1) * 3;  // this wave's row tiles (of 6)
  const int ct0 = (wave & 1) * 3;   // this wave's column tiles (of 6)

  // raw states of this thread's staging tasks (at most 2 x 16 samples: 64
  // genes x K1_KCH / 16 tasks over 256 threads) for chunk k0, loaded one chunk
  // ahead so the loads run under the previous chunk's MFMAs
  static_assert(K1_ROWS * (K1_KCH / 16) <= 2 * 256, "two staging tasks per thread");
  auto fetch = [&](int k0n, uint4 *dst) {
    const int vecsn = ((NP - k0n) < K1_KCH ? (NP - k0n) : K1_KCH) >> 4;
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int v = tid + 256 * j;
      dst[j] = make_uint4(0x03030303u, 0x03030303u, 0x03030303u, 0x03030303u);
      if (v < K1_ROWS * vecsn) {
        const int r = v / vecsn, c = v - r * vecsn;
        const int gene = (r < K1_TB) ? bi * K1_TB + r : bj * K1_TB + (r - K1_TB);
        if (gene < G) dst[j] = *(const uint4 *)(xs + (size_t)gene * NP + k0n + c * 16);
      }
    }
  };
  uint4 nxt[2];
  fetch(0, nxt);
  for (int k0 = 0; k0 < NP; k0 += K1_KCH) {
    const int kch = (NP - k0) < K1_KCH ? (NP - k0) : K1_KCH;
    const uint4 cur[2] = {nxt[0], nxt[1]};
    __syncthreads();
    // stage the one-hot operands of 64 genes x kch samples (genes 0..31 =
    // block bi, split by class; 32..63 = block bj), 16 samples per task
    const int vecs = kch >> 4;
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int v = tid + 256 * j;
      if (v >= K1_ROWS * vecs) continue;
      const int r = v / vecs, c = v - r * vecs;
      const uint4 x = cur[j];
      uint4 h0, h1, h2;
      one_hot3(x.x, h0.x, h1.x, h2.x);
      one_hot3(x.y, h0.y, h1.y, h2.y);
      one_hot3(x.z, h0.z, h1.z, h2.z);
      one_hot3(x.w, h0.w, h1.w, h2.w);
      if (r < K1_TB) {
        uint8_t *d1 = opA1 + (r * 3) * K1_RS + c * 16;
        if (MODE != K1_BOTH) {  // pooled: every sample; class 1: compacted rows
          *(uint4 *)(d1) = h0;
          *(uint4 *)(d1 + K1_RS) = h1;
          *(uint4 *)(d1 + 2 * K1_RS) = h2;
        } else {
          const uint4 m = class1_bytes(k0 + c * 16);
          *(uint4 *)(d1) = make_uint4(h0.x & m.x, h0.y & m.y, h0.z & m.z, h0.w & m.w);
          *(uint4 *)(d1 + K1_RS) = make_uint4(h1.x & m.x, h1.y & m.y, h1.z & m.z, h1.w & m.w);
          *(uint4 *)(d1 + 2 * K1_RS) = make_uint4(h2.x & m.x, h2.y & m.y, h2.z & m.z, h2.w & m.w);
          const uint4 m2 = make_uint4(m.x ^ 0x01010101u, m.y ^ 0x01010101u, m.z ^ 0x01010101u, m.w ^ 0x01010101u);
          uint8_t *d2 = opA2 + (r * 3) * K1_RS + c * 16;
          *(uint4 *)(d2) = make_uint4(h0.x & m2.x, h0.y & m2.y, h0.z & m2.z, h0.w & m2.w);
          *(uint4 *)(d2 + K1_RS) = make_uint4(h1.x & m2.x, h1.y & m2.y, h1.z & m2.z, h1.w & m2.w);
          *(uint4 *)(d2 + 2 * K1_RS) = make_uint4(h2.x & m2.x, h2.y & m2.y, h2.z & m2.z, h2.w & m2.w);
        }
      } else {
        uint8_t *d = opB + ((r - K1_TB) * 3) * K1_RS + c * 16;
        *(uint4 *)(d) = h0;
        *(uint4 *)(d + K1_RS) = h1;
        *(uint4 *)(d + 2 * K1_RS) = h2;
      }
    }
    __syncthreads();
    if (k0 + K1_KCH < NP) fetch(k0 + K1_KCH, nxt);
    for (int kk = 0; kk < kch; kk += 64) {
      const int kofs = kk + (lane >> 4) * 16;
      v4i A1[3], A2[3], B[3];
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const int R = (rt0 + i) * 16 + (lane & 15);
        const int C = (ct0 + i) * 16 + (lane & 15);
        A1[i] = *(const v4i *)(opA1 + R * K1_RS + kofs);
        if (MODE == K1_BOTH) A2[i] = *(const v4i *)(opA2 + R * K1_RS + kofs);
        B[i] = *(const v4i *)(opB + C * K1_RS + kofs);
      }
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
          acc1[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A1[i], B[j], acc1[i][j], 0, 0, 0);
          if (MODE == K1_BOTH) acc2[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A2[i], B[j], acc2[i][j], 0, 0, 0);
        }
    }
  }
  // class 2 = pooled - class 1: the pooled table of this thread's next pair
  // is loaded one pair ahead (the first before the table barrier), so the
  // loads' latency hides under the tests
  auto tot_load = [&](int idx, uint64_t &q0, uint64_t &q1, uint64_t &q2) {
    const int li = idx >> 5, lj = idx & 31, gi = bi * K1_TB + li, gj = bj * K1_TB + lj;
    if (MODE == K1_CLASS1 && idx < K1_TB * K1_TB && gi < G && gj < gi)
      load_tally_words(totals_at(a, sm, pair_index(gi, gj)), a.tabw, a.N, q0, q1, q2);
  };
  uint64_t nq0 = 0, nq1 = 0, nq2 = 0;
  tot_load(tid, nq0, nq1, nq2);
  __syncthreads();
  phase_mark(a, 13, tph);
  // accumulators -> LDS tables T[g][R][C] (u16)
  uint16_t *T = (uint16_t *)rows;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        int R = (rt0 + i) * 16 + 4 * (lane >> 4) + e;
        int C = (ct0 + j) * 16 + (lane & 15);
        T[R * K1_TS + C] = (uint16_t)acc1[i][j][e];
        if (MODE == K1_BOTH) T[96 * K1_TS + R * K1_TS + C] = (uint16_t)acc2[i][j][e];
      }
  __syncthreads();

  const double cpairs = (double)sm.c;
  int ntests = 0;
  // idx = li * 32 + lj: a wave covers rows li, li + 1 of the block, lanes
  // 0..31 and 32..63; its ballots of the full decisions are the two rows'
  // full-edge words of column block bj (the loo kernel's plan reads them)
  for (int idx = tid; idx < K1_TB * K1_TB; idx += 256) {
    const uint64_t q0 = nq0, q1 = nq1, q2 = nq2;
    tot_load(idx + 256, nq0, nq1, nq2);
    int li = idx >> 5, lj = idx & 31;
    int gi = bi * K1_TB + li, gj = bj * K1_TB + lj;
    int f1 = 0, f2 = 0;
    if (gi < G && gj < gi) {
    int t1[9], t2[9];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int k = 0; k < 3; k++) t1[r * 3 + k] = T[(li * 3 + r) * K1_TS + lj * 3 + k];
    long long p = pair_index(gi, gj);
    if (MODE == K1_POOLED) {
      store_tally(a.totals + (size_t)(sm.pairbase + p) * a.tabw, t1, a.tabw);
      continue;  // the loop's ballots below only serve the test modes (fbits is null here)
    }
    if (MODE == K1_BOTH) {
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int k = 0; k < 3; k++) t2[r * 3 + k] = T[96 * K1_TS + (li * 3 + r) * K1_TS + lj * 3 + k];
    } else {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        t2[i] = (int)((q0 >> (16 * i)) & 0xffff) - t1[i];
        t2[4 + i] = (int)((q1 >> (16 * i)) & 0xffff) - t1[4 + i];
      }
      t2[8] = (int)(q2 & 0xffff) - t1[8];
    }
    int pr = (int)((sh_pb[li] >> lj) & 1u);
    {
    uint32_t info = 0;
    const bool logged = im.perm > 0 && im.perm <= a.label_perms;
    if (!a.win || im.perm == 0 || logged || gj <= sh_win[li]) {
      if (MODE == K1_CLASS1) {
        f1 = eddy_edge_screened_only(t1, pr, &sh_thr[0]);
        f2 = eddy_edge_screened_only(t2, pr, &sh_thr[1]);
        if (f1 < 0) undec[atomicAdd(&nundec, 1)] = (uint16_t)(idx << 1);
        if (f2 < 0) undec[atomicAdd(&nundec, 1)] = (uint16_t)(idx << 1 | 1);
        f1 = f1 > 0;
        f2 = f2 > 0;
      } else {
        f1 = eddy_edge_decision_screened(t1, pr, &a.ep, &sh_eb, cpairs, &sh_thr[0]);
        f2 = eddy_edge_decision_screened(t2, pr, &a.ep, &sh_eb, cpairs, &sh_thr[1]);
      }
      ntests += 2;
      info = (f1 ? INFO_FULL(0) : 0u) | (f2 ? INFO_FULL(1) : 0u) | INFO_FEVAL(0) | INFO_FEVAL(1);
      if (a.labhits && im.perm > 0 && !im.nolog && (a.label_perms == 0 || logged))
        atomicAdd(&a.labhits[sm.pairbase + p], (unsigned)(f1 + f2));
    }
    if (MODE == K1_BOTH) {
      info |= eddy_loo_mask_fast(t1, f1, pr, &a.ep, &sh_eb, cpairs, &ntests) << INFO_LOO_SHIFT(0);
      info |= eddy_loo_mask_fast(t2, f2, pr, &a.ep, &sh_eb, cpairs, &ntests) << INFO_LOO_SHIFT(1);
      info |= INFO_EVAL(0) | INFO_EVAL(1);
    } else {
      store_tally(tally_at(a, im, p), t1, a.tabw);
    }
    a.info[im.infooff + p] = info;
    }
    }
    const uint64_t b1 = __ballot(f1), b2 = __ballot(f2);
    if (a.fbits && (lane & 31) == 0 && gi < G) {
      const int h = lane >> 5;  // 0: row li (lanes 0..31), 1: row li + 1
      const uint32_t w1 = (uint32_t)(b1 >> (32 * h)), w2 = (uint32_t)(b2 >> (32 * h));
      a.fbits[(size_t)(im.goff * 2 + gi) * a.fbw + bj] = w1;
      a.fbits[(size_t)(im.goff * 2 + G + gi) * a.fbw + bj] = w2;
    }
  }
  atomicAdd(&tcount, ntests);
  __syncthreads();
  if (MODE == K1_CLASS1) {
    // the open full tests, one lane each: an edge sets its bit in the info
    // word and the full-edge word stored above, and counts for the labels
    const bool logged = im.perm > 0 && im.perm <= a.label_perms;
    const int nq = nundec;
    for (int e = tid; e < nq; e += 256) {
      const int idx = undec[e] >> 1, g = undec[e] & 1;
      const int li = idx >> 5, lj = idx & 31, gi = bi * K1_TB + li, gj = bj * K1_TB + lj;
      int t[9];
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int k = 0; k < 3; k++) t[r * 3 + k] = T[(li * 3 + r) * K1_TS + lj * 3 + k];
      const long long p = pair_index(gi, gj);
      if (g) {
        uint64_t q0, q1, q2;
        load_tally_words(totals_at(a, sm, p), a.tabw, a.N, q0, q1, q2);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          t[i] = (int)((q0 >> (16 * i)) & 0xffff) - t[i];
          t[4 + i] = (int)((q1 >> (16 * i)) & 0xffff) - t[4 + i];
        }
        t[8] = (int)(q2 & 0xffff) - t[8];
      }
      const int pr = (int)((sh_pb[li] >> lj) & 1u);
      if (eddy_edge_decision_fast(t, pr, &a.ep, &sh_eb, cpairs)) {
        atomicOr(&a.info[im.infooff + p], (uint32_t)INFO_FULL(g));
        if (a.fbits) atomicOr(&a.fbits[(size_t)(im.goff * 2 + g * G + gi) * a.fbw + bj], 1u << lj);
        if (a.labhits && im.perm > 0 && !im.nolog && (a.label_perms == 0 || logged))
          atomicAdd(&a.labhits[sm.pairbase + p], 1u);
      }
    }
  }
  if (tid == 0 && a.tests) atomicAdd((unsigned long long *)&a.tests[item], (unsigned long long)tcount);
  phase_mark(a, 14, tph);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_count_test(KArgs a) {
  count_test_body<K1_CLASS1>(a);
}

// The rows of each item's set restricted to the item's class-1 samples, in
// ascending sample order, padded to a.kp with state 3 (one-hot zero): row g
// of item i at a.xc + (goff_i + g) * a.kp.  The cutoff schedule's count
// kernel contracts over these n_A samples instead of all N.  One workgroup per
// item: the class-1 sample list in LDS (a ballot per 64 samples), then each
// thread owns four output positions (their sample ids in registers) and
// walks the rows, four independent byte loads per row (the set's rows are
// L2-resident, shared by its permutations) and one 4-byte store.
__global__ __launch_bounds__(256) void k_gather_class1(KArgs a) {
  const ItemMeta im = a.items[blockIdx.x];
  const SetMeta &sm = a.sets[im.set];
  const int N = a.N, NP = (N + 63) & ~63, KP = a.kp, G = sm.G, NW = NP >> 6;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t *lab = a.labels + im.laboff;
  extern __shared__ __align__(16) unsigned char gdyn[];
  int *base = (int *)gdyn;                   // [NW + 1]: class-1 samples before word w
  uint16_t *c1 = (uint16_t *)(base + NW + 1);  // [KP]: class-1 sample ids
  auto c1word = [&](int w) {
    const int s0 = w * 64;
    uint64_t c2 = (uint64_t)lab[s0 >> 5];
    if (s0 + 32 < N) c2 |= (uint64_t)lab[(s0 >> 5) + 1] << 32;
    const uint64_t valid = N - s0 >= 64 ? ~0ULL : ((1ULL << (N - s0)) - 1ULL);
    return ~c2 & valid;
  };
  for (int w = tid; w < NW; w += 256) base[w + 1] = __popcll(c1word(w));
  __syncthreads();
  if (tid == 0) {
    base[0] = 0;
    for (int w = 1; w <= NW; w++) base[w] += base[w - 1];
  }
  __syncthreads();
  const uint64_t lt = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
  for (int w = wave; w < NW; w += EDDY_NWAVES) {
    const uint64_t m = c1word(w);
    if ((m >> lane) & 1) c1[base[w] + __popcll(m & lt)] = (uint16_t)(w * 64 + lane);
  }
  const int n1 = base[NW];
  for (int k = n1 + tid; k < KP; k += 256) c1[k] = 0xffff;  // padding
  __syncthreads();
  const uint8_t *xs = a.xs + sm.xoff;
  uint32_t *out = (uint32_t *)(a.xc + (size_t)im.goff * KP);
  const int q4 = KP >> 2;
  // row groups: threads tid and tid + q4 take the same positions, other rows
  const int per = q4 <= 256 ? 256 / q4 : 1;
  for (int k4 = tid % q4 + (tid / q4 >= per ? q4 : 0); k4 < q4; k4 += 256) {
    const int grp = q4 <= 256 ? tid / q4 : 0;
    int id[4];
#pragma unroll
    for (int j = 0; j < 4; j++) id[j] = c1[4 * k4 + j];
#pragma unroll 4
    for (int g = grp; g < G; g += per) {
      const uint8_t *row = xs + (size_t)g * NP;
      uint32_t v = 0;
#pragma unroll
      for (int j = 0; j < 4; j++) v |= (uint32_t)(id[j] != 0xffff ? row[id[j]] : 3) << (8 * j);
      out[(size_t)g * q4 + k4] = v;
    }
  }
}

__global__ __launch_bounds__(256) void k_count_test_all(KArgs a) { count_test_body<K1_BOTH>(a); }
__global__ __launch_bounds__(256) void k_count_pooled(KArgs a) { count_test_body<K1_POOLED>(a); }

// Learned-proposal features of every pair of every set from its pooled table
// (all samples: both classes of permutation 0, which no permutation changes),
// into a.feat at the pair's run index.  One block per set, a thread per pair.
__global__ void k_features(KArgs a, int nsets) {
  if ((int)blockIdx.x >= nsets) return;
  const SetMeta &sm = a.sets[blockIdx.x];
  const int zero[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (long long p = threadIdx.x; p < sm.c; p += blockDim.x) {
    uint64_t q0, q1, q2;
    load_tally_words(totals_at(a, sm, p), a.tabw, a.N, q0, q1, q2);
    int t[9];
    unpack_tally(q0, q1, q2, t);
    pair_features(t, zero, prior_bit(a, sm, p), sm.c, a.feat + (sm.pairbase + p) * 12);
  }
}

// --------------------------------------------------------------------------
// K2 (cutoff schedule): leave-one-out tests of each row, class g, up to the
// row's cutoff C = the column of its mp-th "solid" edge (an edge in the full
// test and in every leave-one-out test), or the whole row.  Every network of
// class g takes its first mp parents (kernel3.cu:70-109) from columns <= C,
// so later columns cannot change any capped network.
//
// C is not known before the leave-one-out tests, so the kernel works in
// three phases per item (one workgroup):
//   A  plan: P = column of the (mp + extra)-th full-test edge of the row
//      (every column when the row has fewer); extra = 0 in the fixed plan,
//      or the learned plan's count of full edges predicted not solid
//   B  all (pair, class) entries with column <= P, flattened over the item's
//      rows and tested in parallel (packed: no row-length divergence)
//   C  exact check per row: count solid edges up to P; a row short of mp
//      continues column by column (the fallback) until mp solid edges or the
//      row's end.  Correctness never depends on the plan.
// --------------------------------------------------------------------------
#define K2_MAXROWS (2 * EDDY_MAX_G)

__device__ __forceinline__ uint32_t loo_of_pair(const KArgs &a, const ItemMeta &im, const SetMeta &sm, long long p,
                                                int g, int full, double cpairs, int *ntests) {
  int tt[9];
  load_class_tally(a, im, sm, p, g, tt);
  return d_loo_mask(tt, full, prior_bit(a, sm, p), a, sm, cpairs, ntests);
}

// Class g's 16-bit half of a pair's info word: only the thread handling
// (pair, class) writes it, with plain stores; a __syncthreads separates the
// phases that read what another thread wrote.
__device__ __forceinline__ uint16_t *info_half(const KArgs &a, const ItemMeta &im, long long p, int g) {
  return (uint16_t *)(a.info + im.infooff + p) + g;
}

// full-test decision of (pair, class), evaluated here when the count kernel
// skipped it (outside the learned window or deferred)
__device__ __forceinline__ int full_of(const KArgs &a, const ItemMeta &im, const SetMeta &sm, long long p, int g,
                                       double cpairs, int *ntests) {
  uint16_t *ih = info_half(a, im, p, g);
  const uint16_t h = *ih;
  if (h & IH_FEVAL) return h & IH_FULL;
  int tt[9];
  load_class_tally(a, im, sm, p, g, tt);
  const int f = d_edge_test(tt, prior_bit(a, sm, p), a, sm, cpairs);
  (*ntests)++;
  *ih = (uint16_t)(h | IH_FEVAL | (f ? IH_FULL : 0u));
  return f;
}

// records the leave-one-out mask of (pair, class)
__device__ __forceinline__ void set_loo(const KArgs &a, const ItemMeta &im, long long p, int g, uint32_t m) {
  uint16_t *ih = info_half(a, im, p, g);
  *ih = (uint16_t)(*ih | IH_EVAL | (m << IH_LOO_SHIFT));
}

__device__ int block_exscan(int *v, int n, int *wsum);

__device__ __forceinline__ void loo_body(const KArgs &a) {
  const int item = a.order[blockIdx.x];  // largest items first (shorter tail)
  const ItemMeta im = a.items[item];
  const SetMeta &sm = a.sets[im.set];  // read in place: the bounds are indexed per tally
  const int G = sm.G, mp = a.mp;
  const double cpairs = (double)sm.c;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ int plan[K2_MAXROWS];       // planned last column per (row, class) entry
  __shared__ int off[K2_MAXROWS + 1];    // exclusive prefix of entry counts
  __shared__ int solid[K2_MAXROWS];      // solid edges found up to the plan
  __shared__ int wsum[257];
  __shared__ int tcount, nfb, nfbcols;
  if (tid == 0) tcount = nfb = nfbcols = 0;
  const int R = 2 * G;  // entries r = 2t + g
  int ntests = 0, fb = 0, fbc = 0;
  unsigned long long tph = 0;
  phase_mark(a, -1, tph);
  // With the learned window the count kernel ran the full tests of row t
  // only up to column win[t]; the plan below runs the missing ones it needs
  // (64 columns per wave step).
  // ---- A: plan P = column of the (mp + extra)-th full-test edge, a wave per
  // entry scanning 64 columns per ballot ------------------------------------
  // The count kernel's full-edge words (32 columns of a row and class per
  // word, valid over the columns whose full test it ran: all of them, or
  // those up to the row's window) give the plan without reading the info
  // words: a group of NBWP lanes per entry, one word per lane, a prefix count
  // of their edges, and the want-th edge's bit.  An entry whose evaluated
  // columns hold fewer than `want` full edges but do not cover the row
  // continues below, column by column.
  {
    const int NBW = a.fbw, NBWP = a.fbwp, RPW = 64 / NBWP;
    const int gidx = lane / NBWP, gl = lane - gidx * NBWP;
    const bool logged = im.perm > 0 && im.perm <= a.label_perms;
    const bool allev = !a.win || im.perm == 0 || logged;
    for (int r0 = wave * RPW; r0 < R; r0 += EDDY_NWAVES * RPW) {
      const int r = r0 + gidx;
      const bool act = r < R;
      const int t = act ? (r >> 1) : 0, g = r & 1;
      const int want = act ? mp + (a.plan_extra ? a.plan_extra[(im.goff * 2 + g * G + t)] : 0) : 1;
      const int L = !act ? 0 : allev ? t : min(t, max(0, a.win[sm.genebase + t] + 1));
      const int nw = (L + 31) >> 5;
      uint32_t wd = 0;
      if (gl < nw) {
        wd = a.fbits[(size_t)(im.goff * 2 + g * G + t) * NBW + gl];
        if (gl == nw - 1 && (L & 31)) wd &= (1u << (L & 31)) - 1u;
      }
      const int c = __popc(wd);
      int inc = c;
      for (int d = 1; d < NBWP; d <<= 1) {
        const int y = __shfl_up(inc, d, NBWP);
        if (gl >= d) inc += y;
      }
      const int total = __shfl(inc, NBWP - 1, NBWP);
      if (act && inc >= want && inc - c < want) {
        uint32_t x = wd;
        for (int i = 1; i < want - (inc - c); i++) x &= x - 1;
        const int P = gl * 32 + __ffs(x) - 1;
        plan[r] = P;
        off[r] = P + 1;
        solid[r] = 0;
      }
      if (act && gl == 0 && total < want) {
        if (L >= t) {  // the whole row
          plan[r] = t - 1;
          off[r] = t;
          solid[r] = 0;
        } else {  // continue from column L with `total` full edges so far
          plan[r] = -2 - L;
          solid[r] = total;
        }
      }
    }
  }
  __syncthreads();
  for (int r = wave; r < R; r += EDDY_NWAVES) {
    if (plan[r] > -2) continue;  // wave-uniform
    const int t = r >> 1, g = r & 1;
    const int want = mp + (a.plan_extra ? a.plan_extra[(im.goff * 2 + g * G + t)] : 0);
    const int L = -2 - plan[r];
    int hits = solid[r], P = t - 1;
    for (int base = L; base < t; base += 64) {
      const int col = base + lane;
      int f = 0;
      if (col < t) f = full_of(a, im, sm, pair_index(t, col), g, cpairs, &ntests);
      const uint64_t b = __ballot(f);
      const int nb = __popcll(b);
      if (hits + nb >= want) {  // wave-uniform
        uint64_t x = b;
        for (int i = 1; i < want - hits; i++) x &= x - 1;
        P = base + __ffsll((long long)x) - 1;
        break;
      }
      hits += nb;
    }
    wave_sync();
    if (lane == 0) {
      plan[r] = P;
      off[r] = P + 1;
      solid[r] = 0;
    }
  }
  __syncthreads();
  phase_mark(a, 10, tph);
  const int total = block_exscan(off, R, wsum);
  if (tid == 0) off[R] = total;
  __syncthreads();
  // ---- B: packed leave-one-out tests of every planned (pair, class) ------
  // 64 consecutive entries per wave step: one (wave-uniform) binary search
  // for the step's first entry, then each lane walks to its own row.  The
  // next step's entry is located and its info half, prior word and tally
  // loaded before the current entry's tests, so the loads run under them
  // (an entry's info half is written only by its own lane in this phase).
  struct LooEntry {
    int lo, g, ok;
    long long p;
    uint32_t ih, pw;
    uint64_t q0, q1, q2;
  };
  auto locate = [&](int e0, LooEntry &x) {
    x.ok = 0;
    if (e0 >= total) return;
    int lo = 0, hi = R - 1;  // last r with off[r] <= e0
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (off[mid] <= e0) lo = mid; else hi = mid - 1;
    }
    const int e = e0 + lane;
    if (e >= total) return;
    while (off[lo + 1] <= e) lo++;  // off[R] = total > e ends the walk
    const int t = lo >> 1, g = lo & 1, col = e - off[lo];
    const long long p = pair_index(t, col);
    x.ok = 1;
    x.lo = lo;
    x.g = g;
    x.p = p;
    x.ih = *info_half(a, im, p, g);
    x.pw = a.prior[sm.poff + (p >> 5)];
    class_tally_words(a, im, sm, p, g, x.q0, x.q1, x.q2);
  };
  // Most entries are decided all at once (eddy_loo_mask_try); the others
  // need the per-cell tests, a long path that one such lane would make its
  // whole wave take.  They go to the wave's queue instead (row entry lo,
  // column, full decision), which the wave drains 64 at a time, every lane
  // on one entry.
  __shared__ uint32_t hq[EDDY_NWAVES][128];
  uint32_t *q = hq[wave];
  const uint64_t lt = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
  int hn = 0;  // entries in the queue (the same in every lane)
  auto drain = [&](int cnt) {
    wave_sync();
    if (lane < cnt) {
      const uint32_t w = q[lane];
      const int lo = (int)(w & 0xfffu), col = (int)((w >> 12) & 0x7ffffu), full = (int)(w >> 31);
      const int t = lo >> 1, g = lo & 1;
      const long long p = pair_index(t, col);
      int tt[9];
      load_class_tally(a, im, sm, p, g, tt);
      const uint32_t m = eddy_loo_mask_cells(tt, full, prior_bit(a, sm, p), &a.ep, &sm.eb, cpairs, &ntests);
      *info_half(a, im, p, g) = (uint16_t)(IH_FEVAL | (full ? IH_FULL : 0u) | IH_EVAL | (m << IH_LOO_SHIFT));
      if (full && m == 0x1FFu) atomicAdd(&solid[lo], 1);
    }
    wave_sync();
    // the rest moves to the front
    if (lane < hn - cnt) q[lane] = q[cnt + lane];
    hn -= cnt;
    wave_sync();
  };
  LooEntry nx;
  locate(wave * 64, nx);
  for (int e0 = wave * 64; e0 < total; e0 += 256) {
    const LooEntry cu = nx;
    locate(e0 + 256, nx);
    bool hard = false;
    uint32_t hw = 0;
    if (cu.ok) {
      int tt[9];
      unpack_tally(cu.q0, cu.q1, cu.q2, tt);
      const int pr = (int)((cu.pw >> (cu.p & 31)) & 1u);
      uint32_t h = cu.ih;
      int full;
      if (h & IH_FEVAL) {
        full = (int)(h & IH_FULL);
      } else {  // outside the learned window or deferred: the full test here
        full = d_edge_test(tt, pr, a, sm, cpairs);
        ntests++;
        h |= IH_FEVAL | (full ? IH_FULL : 0u);
      }
      uint32_t m;
      if (eddy_loo_mask_try(tt, full, pr, &a.ep, &sm.eb, &m)) {
        *info_half(a, im, cu.p, cu.g) = (uint16_t)(h | IH_EVAL | (m << IH_LOO_SHIFT));
        if (full && m == 0x1FFu) atomicAdd(&solid[cu.lo], 1);
      } else {
        hard = true;
        const int col = (int)(cu.p - pair_index(cu.lo >> 1, 0));
        hw = (uint32_t)cu.lo | ((uint32_t)col << 12) | ((uint32_t)full << 31);
      }
    }
    const uint64_t hb = __ballot(hard);
    if (hard) q[hn + __popcll(hb & lt)] = hw;
    hn += __popcll(hb);
    if (hn >= 64) drain(64);
  }
  if (hn > 0) drain(hn);
  __syncthreads();
  phase_mark(a, 11, tph);
  // ---- C: a row with fewer than mp solid edges up to its plan continues
  // past it (a wave per row) until mp solid edges or the row's end: the next
  // columns up to the (mp - solid)-th further full-test edge all lie before
  // the cutoff (solid edges are full edges), so they are tested together,
  // 64 per step.  Correctness never depends on the plan. ------------------
  for (int r = wave; r < R; r += EDDY_NWAVES) {
    const int t = r >> 1, g = r & 1;
    int have = solid[r];
    if (have >= mp || plan[r] >= t - 1) continue;  // wave-uniform
    fb += lane == 0;
    int base = plan[r] + 1;
    while (have < mp && base < t) {
      // end of this round: the (mp - have)-th full edge at or after base
      int need = mp - have, P2 = t - 1;
      for (int b0 = base; b0 < t; b0 += 64) {
        const int col = b0 + lane;
        int f = 0;
        if (col < t) {
          const long long pc = pair_index(t, col);
          f = full_of(a, im, sm, pc, g, cpairs, &ntests);
        }
        const uint64_t b = __ballot(f);
        const int nb = __popcll(b);
        if (nb >= need) {
          uint64_t x = b;
          for (int i = 1; i < need; i++) x &= x - 1;
          P2 = b0 + __ffsll((long long)x) - 1;
          break;
        }
        need -= nb;
      }
      for (int b0 = base; b0 <= P2; b0 += 64) {
        const int col = b0 + lane;
        int sol = 0;
        if (col <= P2) {
          const long long p = pair_index(t, col);
          const int full = full_of(a, im, sm, p, g, cpairs, &ntests);
          const uint32_t m = loo_of_pair(a, im, sm, p, g, full, cpairs, &ntests);
          set_loo(a, im, p, g, m);
          sol = full && m == 0x1FFu;
          fbc++;
        }
        have += __popcll(__ballot(sol));
      }
      base = P2 + 1;
    }
  }
  atomicAdd(&tcount, ntests);
  atomicAdd(&nfb, fb);
  atomicAdd(&nfbcols, fbc);
  __syncthreads();
  phase_mark(a, 12, tph);
  if (tid == 0) {
    atomicAdd((unsigned long long *)&a.tests[item], (unsigned long long)tcount);
    if (a.counters) {
      cnt_add(a, 0, (unsigned long long)nfb);      // rows needing fallback
      cnt_add(a, 1, (unsigned long long)nfbcols);  // columns tested in fallback
      cnt_add(a, 2, (unsigned long long)total);    // planned entries
      cnt_add(a, 3, (unsigned long long)R);        // rows
    }
  }
}

// two register budgets of the same code: 4 waves per SIMD (128 VGPRs, a few
// bytes spilled) or 3 (no spills); the launch configuration picks one
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_loo(KArgs a) { loo_body(a); }
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_loo_w3(KArgs a) { loo_body(a); }

// zeroes a slot's per-item counters before its kernels (a kernel rather than
// memsets, so the launch sequence can be one graph)
__global__ void k_clear_slot(int *evcount, int *status, long long *tests, unsigned *bdcount, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    evcount[i] = 0;
    status[i] = 0;
    tests[i] = 0;
  }
  if (i == 0) *bdcount = 0u;
}

// --------------------------------------------------------------------------
// shared per-item helpers
// --------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// hash of row t's parent list par[0..n): begin, one step per parent, end
__device__ __forceinline__ uint64_t row_hash_begin(uint64_t seed, int t) {
  return mix64(seed ^ (0x9E3779B97F4A7C15ULL * (uint64_t)(t + 1)));
}
__device__ __forceinline__ uint64_t row_hash_step(uint64_t h, int par, int i) {
  return mix64(h ^ (((uint64_t)par + 1) << 16) ^ (uint64_t)i);
}
__device__ __forceinline__ uint64_t row_hash_end(uint64_t h, int n) { return mix64(h ^ (uint64_t)n * 0xD6E8FEB86659FD93ULL); }
__device__ __forceinline__ uint64_t row_hash(uint64_t seed, int t, int n, const uint16_t *par) {
  uint64_t h = row_hash_begin(seed, t);
  for (int i = 0; i < n; i++) h = row_hash_step(h, par[i], i);
  return row_hash_end(h, n);
}

// class lists: cls[g][k] = k-th sample of class g in ascending sample order
__device__ void build_class_lists(const uint32_t *lab, int N, uint16_t *cls0, uint16_t *cls1, int *counts) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ int wsum[2][EDDY_NWAVES];
  __shared__ int base[2];
  if (tid == 0) base[0] = base[1] = 0;
  __syncthreads();
  for (int s0 = 0; s0 < N; s0 += 64 * EDDY_NWAVES) {
    int s = s0 + tid;
    int v = s < N;
    int c2 = v ? (lab[s >> 5] >> (s & 31)) & 1 : 0;
    uint64_t b1 = __ballot(v && !c2), b2 = __ballot(v && c2);
    uint64_t lt = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
    int r1 = __popcll(b1 & lt), r2 = __popcll(b2 & lt);
    if (lane == 0) {
      wsum[0][wave] = __popcll(b1);
      wsum[1][wave] = __popcll(b2);
    }
    __syncthreads();
    int o1 = base[0], o2 = base[1];
    for (int w = 0; w < wave; w++) {
      o1 += wsum[0][w];
      o2 += wsum[1][w];
    }
    if (v && !c2) cls0[o1 + r1] = (uint16_t)s;
    if (v && c2) cls1[o2 + r2] = (uint16_t)s;
    __syncthreads();
    if (tid == 0) {
      for (int w = 0; w < EDDY_NWAVES; w++) {
        base[0] += wsum[0][w];
        base[1] += wsum[1][w];
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    counts[0] = base[0];
    counts[1] = base[1];
  }
  __syncthreads();
}

// --------------------------------------------------------------------------
// K3: rows.  Per item: for each row t and class g, the row's cutoff C (the
// column of its mp-th solid edge, or t-1), its capped parent list in the full
// network (kernel3.cu:70-109, 111-132), and the leave-one-out networks whose
// row differs (events).  Network hash = XOR over rows of row_hash.
// One workgroup (4 waves) per item; a wave owns one row at a time.
// --------------------------------------------------------------------------
// GM: the S network deltas and the class lists in global memory (the item's
// hash words and class-list slots) instead of LDS, for sample counts whose
// arrays do not fit; WIDE: parent caps above EDDY_NARROW_MP, where a
// leave-one-out row list is compared with the full one as it is produced and
// written straight to the event record (no per-lane copy).
template <bool GM, bool WIDE>
__device__ __forceinline__ void rows_body(const KArgs &a) {
  const int item = a.order[blockIdx.x];  // largest items first (shorter tail)
  const ItemMeta im = a.items[item];
  const SetMeta sm = a.sets[im.set];
  const int G = sm.G, N = a.N, S = a.S, mp = a.mp, MPW = a.mpw;
  const int NP = (N + 63) & ~63;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint8_t *xs = a.xs + sm.xoff;

  // LDS sized by the launch (rows_lds_bytes): the S network deltas (not GM),
  // per-wave candidate columns of the largest set, per-wave full row lists,
  // the class lists (not GM), state flags
  extern __shared__ __align__(16) unsigned char rdyn[];
  const int GMX = a.rows_maxg;
  unsigned char *lp = rdyn;
  unsigned long long *delta;
  if (GM) {
    delta = (unsigned long long *)(a.hnet + im.soff);
  } else {
    delta = (unsigned long long *)lp;
    lp += (size_t)S * 8;
  }
  uint32_t *candb = (uint32_t *)lp;                                        // [4][GMX]: col | flip << 16 | full << 25
  lp += (size_t)EDDY_NWAVES * GMX * 4;
  uint16_t *lrowb = (uint16_t *)lp;                                        // [4][MPW]
  lp += (size_t)EDDY_NWAVES * MPW * 2;
  uint16_t *clsb;                                                          // [N]: class 1, then class 2
  if (GM) {
    clsb = a.clsg + im.soff;
  } else {
    clsb = (uint16_t *)lp;
    lp += (size_t)N * 2;
  }
  uint8_t *sflb = (uint8_t *)lp;                                           // [2][GMX]
  uint16_t *cls0 = clsb, *cls1 = clsb + a.n1;
  uint32_t *cand_w = candb + wave * GMX;
  // the item's events are numbered in LDS (this workgroup is the item's only
  // writer) and the count stored once at the end
  __shared__ int sh_evn;
  if (tid == 0) sh_evn = 0;
  uint16_t *lrow = lrowb + wave * MPW;
  __shared__ int ncls[2];
  __shared__ unsigned long long hfull[2];
  __shared__ unsigned long long rchg[EDDY_NWAVES];  // per wave: samples a row's flips touch
  __shared__ unsigned long long rtog[EDDY_NWAVES][128];  // per wave: each candidate's toggled samples

  unsigned long long tph = 0;
  phase_mark(a, -1, tph);
  build_class_lists(a.labels + im.laboff, N, cls0, cls1, ncls);
  if (!GM) {
    for (int k = tid; k < ncls[0]; k += 256) a.clsg[im.soff + k] = cls0[k];
    for (int k = tid; k < ncls[1]; k += 256) a.clsg[im.soff + a.n1 + k] = cls1[k];
  }
  for (int s = tid; s < S; s += 256) delta[s] = 0ULL;
  for (int i = tid; i < 2 * GMX; i += 256) sflb[i] = 0;
  if (tid == 0) hfull[0] = hfull[1] = 0ULL;
  __syncthreads();

  // observed states per gene per class on the class's full data
  // (kernel1.cu:338-387), and the class's state masks: per class and chunk of
  // 8 mask words, each lane holds its samples' ids (the class list is the
  // same for every gene), then a wave per gene gathers the 8 state bytes
  // together and takes a ballot per state
  for (int g = 0; g < 2; g++) {
    const int ng = ncls[g];
    const uint16_t *cl = g ? cls1 : cls0;
    for (int w0 = 0; w0 < a.bitw; w0 += 8) {
      int sid[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int k = (w0 + u) * 64 + lane;
        sid[u] = k < ng ? (int)cl[k] : -1;
      }
      // lane u < 8 stores mask word w0 + u of each state
      const int nw = min(8, a.bitw - w0);
      for (int gene = wave; gene < G; gene += EDDY_NWAVES) {
        const uint8_t *row = xs + (size_t)gene * NP;
        int v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = sid[u] >= 0 ? row[sid[u]] : 3;
        uint64_t *mk = a.bits + ((size_t)(im.goff * 2 + g * G + gene) * 3) * a.bitw;
        uint64_t any0 = 0, any1 = 0, any2 = 0, s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const uint64_t m0 = __ballot(v[u] == 0), m1 = __ballot(v[u] == 1), m2 = __ballot(v[u] == 2);
          if (lane == u) {
            s0 = m0;
            s1 = m1;
            s2 = m2;
          }
          if (u < nw) {
            any0 |= m0;
            any1 |= m1;
            any2 |= m2;
          }
        }
        if (lane < nw) {
          mk[w0 + lane] = s0;
          mk[a.bitw + w0 + lane] = s1;
          mk[2 * a.bitw + w0 + lane] = s2;
        }
        // this wave owns the gene in every chunk: a plain read-modify-write
        if (lane == 0)
          sflb[g * GMX + gene] |= (uint8_t)((any0 ? 1u : 0u) | (any1 ? 2u : 0u) | (any2 ? 4u : 0u));
      }
    }
  }
  __syncthreads();
  for (int i = tid; i < 2 * G; i += 256) {
    const int g = i / G, gene = i - g * G;
    a.sflags[im.goff * 2 + i] = sflb[g * GMX + gene];
  }
  __syncthreads();

  phase_mark(a, 7, tph);
  uint64_t lt_mask = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
  for (int rg = wave; rg < 2 * G; rg += EDDY_NWAVES) {
    const int t = rg >> 1, g = rg & 1;
    // candidates: the columns up to the cutoff that are a full edge or whose
    // leave-one-out decisions differ from the full one (flip mask); a
    // leave-one-out network's row list is drawn from these alone
    int nL = 0, nSolid = 0, nflip = 0, ncand = 0;
    bool done = (t == 0);
    for (int base = 0; base < t && !done; base += 64) {
      int col = base + lane;
      bool valid = col < t;
      uint32_t w = valid ? a.info[im.infooff + pair_index(t, col)] : 0u;
      int full = INFO_FULL_BIT(w, g);
      uint32_t loo = (w >> INFO_LOO_SHIFT(g)) & 0x1FFu;
      bool ev = (w & INFO_EVAL(g)) != 0 && (w & INFO_FEVAL(g)) != 0;  // both tests of this cell ran
      bool solid = valid && full && ev && loo == 0x1FFu;
      uint64_t sm_ = __ballot(solid);
      int need = mp - nSolid;
      int limit = 63;
      if (__popcll(sm_) >= need) {
        uint64_t x = sm_;
        for (int i = 1; i < need; i++) x &= x - 1;
        limit = __ffsll((long long)x) - 1;
        done = true;
      }
      bool inRange = valid && lane <= limit;
      if (__any(inRange && !ev)) {
        if (lane == 0) atomicOr(&a.status[item], ST_NOT_EVALUATED);
      }
      uint64_t hm = __ballot(inRange && full);
      int rank = __popcll(hm & lt_mask);
      if (inRange && full && nL + rank < mp) lrow[nL + rank] = (uint16_t)col;
      nL += __popcll(hm);
      if (nL > mp) nL = mp;
      uint32_t flip = loo ^ (full ? 0x1FFu : 0u);
      const bool isc = inRange && (full || flip != 0);
      const uint64_t cm = __ballot(isc);
      if (isc) cand_w[ncand + __popcll(cm & lt_mask)] = (uint32_t)col | (flip << 16) | ((uint32_t)full << 25);
      ncand += __popcll(cm);
      nflip += __popcll(__ballot(inRange && flip != 0));
      nSolid += __popcll(sm_ & (limit == 63 ? ~0ULL : ((2ULL << limit) - 1)));
      wave_sync();
    }
    wave_sync();
    // full row list
    const size_t lbase = (size_t)(im.goff * 2 + g * G + t);
    if (!WIDE) {
      if (lane < nL) a.lfull[lbase * MPW + lane] = lrow[lane];
    } else {
      for (int i = lane; i < nL; i += 64) a.lfull[lbase * MPW + i] = lrow[i];
    }
    if (lane == 0) a.lcnt[lbase] = (uint16_t)nL;
    uint64_t hrow = row_hash(a.hseed, t, nL, lrow);
    if (lane == 0) atomicXor(&hfull[g], (unsigned long long)hrow);
    if (nflip == 0) {
      wave_sync();
      continue;
    }
    const int ng = ncls[g];
    if (!WIDE && ncand <= 128) {
      // Leave-one-out networks of class g whose row t differs, 64 samples
      // at a time from the class's state masks: for candidate column c with
      // flip mask fm, T_c = OR over the cells (a, b) of fm of (row t's
      // state-a mask AND column c's state-b mask) marks the samples for
      // which the candidate's decision differs from the full one; a sample
      // none marks keeps the full row.  Lane j holds candidates j and 64 + j.
      const uint64_t *mt = a.bits + ((size_t)(im.goff * 2 + g * G + t) * 3) * a.bitw;
      uint32_t ce0 = lane < ncand ? cand_w[lane] : 0u, ce1 = 64 + lane < ncand ? cand_w[64 + lane] : 0u;
      const uint64_t *mc0 = a.bits + ((size_t)(im.goff * 2 + g * G + (int)(ce0 & 0xffffu)) * 3) * a.bitw;
      const uint64_t *mc1 = a.bits + ((size_t)(im.goff * 2 + g * G + (int)(ce1 & 0xffffu)) * 3) * a.bitw;
      const uint32_t fm0 = (ce0 >> 16) & 0x1FFu, fm1 = (ce1 >> 16) & 0x1FFu;
      auto toggles = [&](uint32_t fm, const uint64_t *mc, int w, uint64_t t0, uint64_t t1, uint64_t t2) -> uint64_t {
        if (!fm) return 0ULL;
        const uint64_t c0 = mc[w], c1 = mc[a.bitw + w], c2 = mc[2 * a.bitw + w];
        uint64_t x = 0ULL;
#pragma unroll
        for (int r = 0; r < 3; r++) {
          const uint64_t tr = r == 0 ? t0 : r == 1 ? t1 : t2;
          if ((fm >> (r * 3)) & 1u) x |= tr & c0;
          if ((fm >> (r * 3 + 1)) & 1u) x |= tr & c1;
          if ((fm >> (r * 3 + 2)) & 1u) x |= tr & c2;
        }
        return x;
      };
      for (int w = 0; w * 64 < ng; w++) {
        const uint64_t t0 = mt[w], t1 = mt[a.bitw + w], t2 = mt[2 * a.bitw + w];
        const uint64_t T0 = toggles(fm0, mc0, w, t0, t1, t2), T1 = toggles(fm1, mc1, w, t0, t1, t2);
        if (lane == 0) rchg[wave] = 0ULL;
        unsigned long long *tog = rtog[wave];
        tog[lane] = T0;
        tog[64 + lane] = T1;
        wave_sync();
        if (T0 | T1) atomicOr(&rchg[wave], T0 | T1);
        wave_sync();
        // lane kk takes sample w * 64 + kk when a flip touches it: its row
        // list is the first mp candidates, in column order, whose decision
        // for it (full, toggled by bit kk of the candidate's T) is an edge;
        // compared with the full row and hashed as it is read
        const bool mine = (rchg[wave] >> lane) & 1ULL;
        wave_sync();
        if (!__any(mine)) continue;
        uint64_t h = row_hash_begin(a.hseed, t);
        int cnt = 0;
        bool same = true;
        for (int c = 0; c < ncand; c++) {
          const uint32_t ce = cand_w[c];
          const int bit = (int)((ce >> 25) & 1u) ^ (int)((tog[c] >> lane) & 1ULL);
          if (mine && bit && cnt < mp) {
            const uint16_t col = (uint16_t)(ce & 0xffffu);
            same = same && cnt < nL && lrow[cnt] == col;
            h = row_hash_step(h, col, cnt);
            cnt++;
          }
          if (!__any(mine && cnt < mp)) break;
        }
        const bool ev = mine && !(same && cnt == nL);
        if (!__any(ev)) continue;
        int e = -1;
        uint16_t *pp = nullptr;
        if (ev) {
          const int k = w * 64 + lane;
          const int s = g == 0 ? 1 + k : a.n1 + 1 + k;
          e = atomicAdd(&sh_evn, 1);
          if (e < im.evcap) {
            a.evkey[im.evoff + e] = (uint32_t)s | ((uint32_t)t << 16);
            pp = a.evpar + (size_t)(im.evoff + e) * a.evw;
          } else {
            atomicOr(&a.status[item], ST_EV_OVERFLOW);
          }
          atomicXor(&delta[s], (unsigned long long)(hrow ^ row_hash_end(h, cnt)));
        }
        // the event's parents: the same walk again
        if (__any(pp != nullptr)) {
          uint64_t v = (uint64_t)cnt;
          int c2 = 0;
          for (int c = 0; c < ncand; c++) {
            const uint32_t ce = cand_w[c];
            const int bit = (int)((ce >> 25) & 1u) ^ (int)((tog[c] >> lane) & 1ULL);
            if (pp && bit && c2 < cnt) {
              const uint16_t col = (uint16_t)(ce & 0xffffu);
              if (a.evw == 4) v |= (uint64_t)col << (16 * (c2 + 1));
              else pp[1 + c2] = col;
              c2++;
            }
            if (!__any(pp != nullptr && c2 < cnt)) break;
          }
          if (pp) {
            if (a.evw == 4) *(uint64_t *)pp = v;  // mp = 3: n and three parents in one 8-byte store
            else pp[0] = (uint16_t)cnt;
          }
        }
      }
      wave_sync();
      continue;
    }
    // leave-one-out networks of class g whose row t differs
    const uint8_t *xt = xs + (size_t)t * NP;
    for (int k = lane; k < ng; k += 64) {
      int sid = (g ? cls1 : cls0)[k];
      int vt = xt[sid];
      if (vt >= 3) continue;
      bool changed = false;
      for (int i = 0; i < ncand && !changed; i++) {
        const uint32_t e = cand_w[i], fm = (e >> 16) & 0x1FFu;
        if (!fm) continue;
        const int vc = xs[(size_t)(e & 0xffff) * NP + sid];
        if (vc < 3 && (fm >> (vt * 3 + vc)) & 1) changed = true;
      }
      if (!changed) continue;
      // the sample's row list: the candidates in column order, each a full
      // edge toggled where its flip mask covers the sample's (vt, vc) cell
      auto cand_bit = [cand_w, xs, NP, sid, vt](int i, int &col) -> int {
        const uint32_t e = cand_w[i], fm = (e >> 16) & 0x1FFu;
        col = (int)(e & 0xffff);
        int bit = (int)((e >> 25) & 1u);
        if (fm) {
          const int vc = xs[(size_t)col * NP + sid];
          if (vc < 3) bit ^= (fm >> (vt * 3 + vc)) & 1;
        }
        return bit;
      };
      const int s = g == 0 ? 1 + k : a.n1 + 1 + k;
      if (!WIDE) {
        uint16_t nl[EDDY_NARROW_MP];
        int cnt = 0;
        for (int i = 0; i < ncand && cnt < mp; i++) {
          int col;
          if (cand_bit(i, col)) nl[cnt++] = (uint16_t)col;
        }
        bool same = cnt == nL;
        for (int i = 0; i < cnt && same; i++) same = nl[i] == lrow[i];
        if (same) continue;
        int e = atomicAdd(&sh_evn, 1);
        if (e < im.evcap) {
          a.evkey[im.evoff + e] = (uint32_t)s | ((uint32_t)t << 16);
          uint16_t *pp = a.evpar + (size_t)(im.evoff + e) * a.evw;
          if (a.evw == 4) {  // mp = 3: n and three parents in one 8-byte store
            uint64_t v = (uint64_t)cnt;
            for (int i = 0; i < 3; i++)
              if (i < cnt) v |= (uint64_t)nl[i] << (16 * (i + 1));
            *(uint64_t *)pp = v;
          } else {
            pp[0] = (uint16_t)cnt;
            for (int i = 0; i < cnt; i++) pp[1 + i] = nl[i];
          }
        } else {
          atomicOr(&a.status[item], ST_EV_OVERFLOW);
        }
        atomicXor(&delta[s], (unsigned long long)(hrow ^ row_hash(a.hseed, t, cnt, nl)));
      } else {
        // pass 1: compare with the full row list as the list is produced
        int cnt = 0;
        bool same = true;
        for (int i = 0; i < ncand && cnt < mp; i++) {
          int col;
          if (cand_bit(i, col)) {
            same = same && cnt < nL && lrow[cnt] == (uint16_t)col;
            cnt++;
          }
        }
        if (same && cnt == nL) continue;
        // pass 2: the event record and the row's hash
        int e = atomicAdd(&sh_evn, 1);
        uint16_t *pp = nullptr;
        if (e < im.evcap) {
          a.evkey[im.evoff + e] = (uint32_t)s | ((uint32_t)t << 16);
          pp = a.evpar + (size_t)(im.evoff + e) * a.evw;
          pp[0] = (uint16_t)cnt;
        } else {
          atomicOr(&a.status[item], ST_EV_OVERFLOW);
        }
        uint64_t h = row_hash_begin(a.hseed, t);
        int c2 = 0;
        for (int i = 0; i < ncand && c2 < mp; i++) {
          int col;
          if (cand_bit(i, col)) {
            if (pp) pp[1 + c2] = (uint16_t)col;
            h = row_hash_step(h, col, c2);
            c2++;
          }
        }
        atomicXor(&delta[s], (unsigned long long)(hrow ^ row_hash_end(h, cnt)));
      }
    }
    wave_sync();
  }
  __syncthreads();
  if (tid == 0) a.evcount[item] = sh_evn;
  for (int s = tid; s < S; s += 256) {
    int g = (s <= a.n1) ? 0 : 1;
    if (GM) a.hnet[im.soff + s] ^= hfull[g];  // delta[s] is that word
    else a.hnet[im.soff + s] = hfull[g] ^ delta[s];
  }
  phase_mark(a, 8, tph);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_rows(KArgs a) { rows_body<false, false>(a); }
__global__ __launch_bounds__(256) void k_rows_g(KArgs a) { rows_body<true, false>(a); }
__global__ __launch_bounds__(256) void k_rows_w(KArgs a) { rows_body<false, true>(a); }
__global__ __launch_bounds__(256) void k_rows_gw(KArgs a) { rows_body<true, true>(a); }

// --------------------------------------------------------------------------
// K4: dedup, BDeu, likelihood, JS.  One workgroup per item.



// --------------------------------------------------------------------------

__device__ __forceinline__ int net_class(int s, int n1) { return s <= n1 ? 0 : 1; }

// c / d for 0 <= c < 128 and d in {1, 2, 3} (43/128 = 1/3 + 1/384 is
// exact over that range)
__device__ __forceinline__ int div_small(int c, int d) { return d == 1 ? c : d == 2 ? c >> 1 : (c * 43) >> 7; }

// the c-th (0-based) state present in the 3-bit state set f
__device__ __forceinline__ int nth_state(uint32_t f, int c) {
  return c == 0 ? __builtin_ctz(f) : c == 1 ? __builtin_ctz(f & (f - 1)) : 2;
}

// parent list of row t in network s (events of s sorted by row: es[0..ne),
// their rows et[0..ne))
__device__ __forceinline__ int row_list(const KArgs &a, const ItemMeta &im, int G, int s, int t, const int *es,
                                        const uint16_t *et, int ne, int *ep, const uint16_t **out) {
  while (*ep < ne && et[*ep] < t) (*ep)++;
  if (*ep < ne && et[*ep] == t) {
    const uint16_t *pp = a.evpar + (size_t)(im.evoff + es[*ep]) * a.evw;
    *out = pp + 1;
    return pp[0];
  }
  int g = net_class(s, a.n1);
  size_t lb = (size_t)(im.goff * 2 + g * G + t);
  *out = a.lfull + lb * a.mpw;
  return a.lcnt[lb];
}

__device__ __forceinline__ uint64_t fam_key(int t, int n, const uint16_t *par) {
  if (n > 3) n = 3;  // BDeu uses at most 3 parents (kernel5.cu:119-130)
  uint64_t k = (uint64_t)t | ((uint64_t)n << 10);
  for (int i = 0; i < n; i++) k |= (uint64_t)par[i] << (12 + 10 * i);
  return k + 1;  // 0 marks an empty slot
}

__device__ int ht_insert_min(unsigned long long *keys, int *vals, uint64_t key, int val, int hs) {
  uint32_t h = (uint32_t)mix64(key) & (hs - 1);
  for (int probe = 0; probe < hs; probe++) {
    unsigned long long old = atomicCAS(&keys[h], 0ULL, (unsigned long long)key);
    if (old == 0ULL || old == key) {
      atomicMin(&vals[h], val);
      return (int)h;
    }
    h = (h + 1) & (hs - 1);
  }
  return -1;
}

__device__ int ht_find(const unsigned long long *keys, uint64_t key, int hs) {
  uint32_t h = (uint32_t)mix64(key) & (hs - 1);
  for (int probe = 0; probe < hs; probe++) {
    unsigned long long k = keys[h];
    if (k == key) return (int)h;
    if (k == 0ULL) return -1;
    h = (h + 1) & (hs - 1);
  }
  return -1;
}

// block-wide exclusive scan of n ints in LDS (in place); returns the total
__device__ int block_exscan(int *v, int n, int *wsum) {
  const int tid = threadIdx.x;
  const int per = (n + 255) / 256;
  int b = tid * per, e = min(n, b + per), acc = 0;
  for (int i = b; i < e; i++) acc += v[i];
  wsum[tid] = acc;
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int i = 0; i < 256; i++) {
      int x = wsum[i];
      wsum[i] = run;
      run += x;
    }
    wsum[256] = run;
  }
  __syncthreads();
  acc = wsum[tid];
  for (int i = b; i < e; i++) {
    int x = v[i];
    v[i] = acc;
    acc += x;
  }
  int total = wsum[256];
  __syncthreads();
  return total;
}

// BDeu of one family with two or three parents in one scoring class
// (kernel5.cu:4-236), by a group of GS lanes of one wave (gl = lane within the
// group): the family's state masks staged in the group's LDS, one histogram
// cell per lane, each parent configuration's inner terms per lane, and the
// group's first lane adds the configurations in order (j outer, k inner).
// Every lane of the wave calls it together (it synchronises the wave); a
// group with no family (f < 0) or a family with fewer than two parents only
// keeps step.
__device__ __forceinline__ void bdeu_group(const KArgs &a, const ItemMeta &im, int G, int f, int g, int gl, int GS,
                                           uint64_t *gmst, uint16_t *ghist, double *glgv,
                                           unsigned long long &st_hist, unsigned long long &st_bins) {
  const int W = a.bitw;
  int np = 0, t = 0, p0 = 0, p1 = 0, p2 = 0;
  if (f >= 0) {
    const uint64_t key = a.famkey[im.foff + f] - 1;
    t = (int)(key & 1023);
    np = (int)((key >> 10) & 3);
    p0 = (int)((key >> 12) & 1023);
    p1 = (int)((key >> 22) & 1023);
    p2 = (int)((key >> 32) & 1023);
  }
  const bool mine = np > 1;  // group-uniform
  uint32_t f0 = 1u, f1 = 1u, f2 = 1u, f3 = 1u;
  if (mine) {
    const uint8_t *sf = a.sflags + im.goff * 2 + g * G;
    f0 = sf[t];
    f1 = sf[p0];
    f2 = sf[p1];
    f3 = np > 2 ? sf[p2] : 1u;
  }
  const int r = __popc(f0), n1s = __popc(f1), n2s = __popc(f2), n3s = __popc(f3);
  const int q = n1s * n2s * n3s;  // parents' observed-state combinations
  const bool live = mine && r > 0 && q > 0;
  if (live) {
    const uint64_t *mk = a.bits + (size_t)(im.goff * 2 + g * G) * 3 * W;
    for (int x = gl; x < (np + 1) * 3 * W; x += GS) {
      const int slot = x / W, w = x - slot * W, gix = slot / 3, v = slot - gix * 3;
      const int gene = gix == 0 ? t : gix == 1 ? p0 : gix == 2 ? p1 : p2;
      gmst[x] = mk[(size_t)(gene * 3 + v) * W + w];
    }
  }
  wave_sync();
  if (live) {
    for (int c = gl; c < q * r; c += GS) {
      const int j = div_small(c, r), kk = c - j * r;
      const int j1 = div_small(j, n1s), i1 = j - j1 * n1s;
      const int i3 = div_small(j1, n2s), i2 = j1 - i3 * n2s;
      const uint64_t *m0 = gmst + (0 * 3 + nth_state(f0, kk)) * W;
      const uint64_t *m1 = gmst + (1 * 3 + nth_state(f1, i1)) * W;
      const uint64_t *m2 = gmst + (2 * 3 + nth_state(f2, i2)) * W;
      const uint64_t *m3 = gmst + (3 * 3 + nth_state(f3, i3)) * W;
      int cnt = 0;
      for (int w = 0; w < W; w++) {
        uint64_t x = m0[w] & m1[w] & m2[w];
        if (np > 2) x &= m3[w];
        cnt += __popcll(x);
      }
      ghist[c] = (uint16_t)cnt;
    }
  }
  wave_sync();
  if (live) {
    const double *lq = a.lgtab + (size_t)q * a.lgstride;
    const double *lqr = a.lgtab + (size_t)(q * r) * a.lgstride;
    const double lq0 = lq[0], lqr0 = lqr[0];
    for (int j = gl; j < q; j += GS) {
      int nij = 0;
      double sum1 = 0.0;
      for (int k = 0; k < r; k++) {
        const int h = ghist[j * r + k];
        nij += h;
        sum1 = sum1 + (lqr[h] - lqr0);
      }
      glgv[j] = (lq0 - lq[nij]) + sum1;
    }
  }
  wave_sync();
  if (mine && gl == 0) {
    double sum2 = 0.0;
    if (live)
      for (int j = 0; j < q; j++) sum2 = sum2 + glgv[j];
    a.famscore[(im.foff + f) * 2 + g] = sum2;
    st_hist += (unsigned long long)(1 + np) * (unsigned long long)(g ? a.n2 : a.n1);
    if (live) st_bins += (unsigned long long)(4 * q * r);
  }
  wave_sync();
}

// the BDeu group stage of one wave in an LDS region laid out as
// [waves x G4 groups][12 W] masks, [waves x G4][84] u16 histogram cells,
// [waves x G4][28] configuration terms
__device__ __forceinline__ void bdeu_stage(unsigned char *base, int G4, int W, int wave, int gi, uint64_t **gmst,
                                           uint16_t **ghist, double **glgv) {
  *gmst = (uint64_t *)base + (size_t)(wave * G4 + gi) * 12 * W;
  *ghist = (uint16_t *)((uint64_t *)base + (size_t)EDDY_NWAVES * G4 * 12 * W) + (wave * G4 + gi) * 84;
  *glgv = (double *)((uint64_t *)base + (size_t)EDDY_NWAVES * G4 * 12 * W + ((size_t)EDDY_NWAVES * G4 * 84 * 2 + 7) / 8) +
          (wave * G4 + gi) * 28;
}

// GM: the network hash table and the S-sized network tables in the item's
// global scratch (a.gscr) instead of LDS, for sample counts whose tables do
// not fit
template <bool GM>
__device__ __forceinline__ void score_body(const KArgs &a) {
  const int item = a.order[blockIdx.x];  // largest items first (shorter tail)
  const ItemMeta im = a.items[item];
  const SetMeta sm = a.sets[im.set];
  const int G = sm.G, S = a.S, n1 = a.n1, MPW = a.mpw;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int HT = a.ht_size;

  extern __shared__ __align__(16) unsigned char dyn[];
  // a region shared by phases (the network hash table in phase 1 unless GM,
  // the BDeu mask stage in phase 4), then the network tables (unless GM) and
  // the full networks' families
  unsigned char *gs = GM ? a.gscr + (size_t)item * a.gscr_item : nullptr;
  unsigned long long *hkeys = GM ? (unsigned long long *)gs : (unsigned long long *)dyn;  // [HT]
  int *hvals = (int *)(hkeys + HT);                                                         // [HT]
  int *rep = GM ? hvals + HT : (int *)(dyn + a.score_shared);                               // [S]
  int *evst = rep + S;                                    // [S + 1]
  int *work = evst + S + 1;                               // [S]: scratch counts
  int *ulist = work + S;                                  // [S]: unique networks in order
  int *famfull = GM ? (int *)(dyn + a.score_shared) : ulist + S;  // [2G]
  __shared__ int wsum[257];
  __shared__ int U, F, flag;

  const int E = min(a.evcount[item], im.evcap);
  int *es = a.evorder + im.evoff;
  uint16_t *et = a.evt + im.evoff;              // rows of es[], in the same order
  const uint32_t *ek = a.evkey + im.evoff;      // s | t << 16 per event
  const uint16_t *epar = a.evpar + (size_t)im.evoff * a.evw;  // n, parents per event
  unsigned long long tph = 0;
  phase_mark(a, -1, tph);

  // ---- 1. first occurrence of each network hash -------------------------
  for (int i = tid; i < HT; i += 256) {
    hkeys[i] = 0ULL;
    hvals[i] = 0x7fffffff;
  }
  for (int s = tid; s < S; s += 256) work[s] = 0;
  if (tid == 0) flag = 0;
  __syncthreads();
  for (int s = tid; s < S; s += 256) {
    uint64_t h = a.hnet[im.soff + s];
    if (h == 0ULL) h = 1ULL;
    if (ht_insert_min(hkeys, hvals, h, s, HT) < 0) atomicOr(&flag, 1);
  }
  for (int e = tid; e < E; e += 256) atomicAdd(&work[ek[e] & 0xffffu], 1);
  __syncthreads();
  for (int s = tid; s < S; s += 256) {
    uint64_t h = a.hnet[im.soff + s];
    if (h == 0ULL) h = 1ULL;
    int slot = ht_find(hkeys, h, HT);
    rep[s] = slot >= 0 ? hvals[slot] : s;
    if (slot < 0) atomicOr(&flag, 1);
    evst[s] = work[s];
  }
  __syncthreads();
  // events grouped by network
  block_exscan(evst, S, wsum);
  if (tid == 0) evst[S] = E;
  for (int s = tid; s < S; s += 256) work[s] = 0;
  __syncthreads();
  for (int e = tid; e < E; e += 256) {
    const uint32_t k = ek[e];
    const int s = (int)(k & 0xffffu);
    const int pos = evst[s] + atomicAdd(&work[s], 1);
    es[pos] = e;
    et[pos] = (uint16_t)(k >> 16);
  }
  __syncthreads();
  // sort each network's events by row (a network has at most one event per
  // row, so the order does not depend on the atomics above)
  for (int s = tid; s < S; s += 256) {
    int b = evst[s], n = evst[s + 1] - b;
    for (int i = 1; i < n; i++) {
      const int v = es[b + i];
      const uint16_t tv = et[b + i];
      int j = i - 1;
      while (j >= 0 && et[b + j] > tv) {
        es[b + j + 1] = es[b + j];
        et[b + j + 1] = et[b + j];
        j--;
      }
      es[b + j + 1] = v;
      et[b + j + 1] = tv;
    }
  }
  __syncthreads();
  phase_mark(a, 0, tph);
  // ---- 2. exact confirmation of every merge -----------------------------
  // rows where the two classes' full networks differ, ascending (famfull is
  // free until the families are numbered)
  int *drows = famfull, *dflag = famfull + G;
  __shared__ int ndiff;
  for (int t = tid; t < G; t += 256) {
    const size_t l0 = (size_t)(im.goff * 2 + t), l1 = l0 + G;
    const int c0 = a.lcnt[l0];
    bool d = c0 != a.lcnt[l1];
    for (int i = 0; i < c0 && !d; i++) d = a.lfull[l0 * MPW + i] != a.lfull[l1 * MPW + i];
    dflag[t] = d ? 1 : 0;
  }
  if (tid == 0) ndiff = 0;
  __syncthreads();
  {
    // compact in rounds of 256 rows (ballot per wave, wave totals in wsum)
    for (int t0 = 0; t0 < G; t0 += 256) {
      const int t = t0 + tid;
      const int d = t < G ? dflag[t] : 0;
      const uint64_t m = __ballot(d);
      if (lane == 0) wsum[wave] = __popcll(m);
      __syncthreads();
      int base = ndiff;
      for (int w = 0; w < wave; w++) base += wsum[w];
      if (d) drows[base + __popcll(m & ((1ULL << lane) - 1))] = t;
      __syncthreads();
      if (tid == 0) ndiff += wsum[0] + wsum[1] + wsum[2] + wsum[3];
      __syncthreads();
    }
  }
  // a merge is confirmed row by row over the rows where the two networks can
  // differ: their event rows, and, across classes, the rows of drows
  for (int s = tid; s < S; s += 256) {
    int r = rep[s];
    if (r == s) continue;
    int gs = net_class(s, n1), gr = net_class(r, n1);
    int ns = evst[s + 1] - evst[s], nr = evst[r + 1] - evst[r];
    if (gs == gr && ns == 0 && nr == 0) continue;
    const int nd = gs == gr ? 0 : ndiff;
    bool eq = true;
    int ps = 0, pr = 0, ia = 0, ib = 0, id = 0;
    const int *esS = es + evst[s], *esR = es + evst[r];
    const uint16_t *etS = et + evst[s], *etR = et + evst[r];
    if (gs == gr && ns == nr) {
      // the same class's full network changed at the same rows to the same
      // parent lists: the same network (anything else takes the row walk).
      // Four events at a time, every load of a round issued before its
      // compares; with four u16 per event (mp = 3) a list is one 8-byte word
      // compared up to its count
      bool same = true;
      for (int i0 = 0; i0 < ns && same; i0 += 4) {
        int eS[4], eR[4];
        bool rows_eq = true;
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int i = min(i0 + u, ns - 1);
          eS[u] = esS[i];
          eR[u] = esR[i];
          rows_eq = rows_eq && etS[i] == etR[i];
        }
        if (a.evw == 4) {
          uint64_t vS[4], vR[4];
#pragma unroll
          for (int u = 0; u < 4; u++) {
            vS[u] = *(const uint64_t *)(a.evpar + (size_t)(im.evoff + eS[u]) * 4);
            vR[u] = *(const uint64_t *)(a.evpar + (size_t)(im.evoff + eR[u]) * 4);
          }
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const int c = (int)(vS[u] & 0xffffu);
            const uint64_t m = c >= 3 ? ~0ULL : (1ULL << (16 * (c + 1))) - 1ULL;
            same = same && ((vS[u] ^ vR[u]) & m) == 0ULL;
          }
        } else {
          for (int u = 0; u < 4 && same; u++) {
            const uint16_t *ls = a.evpar + (size_t)(im.evoff + eS[u]) * a.evw;
            const uint16_t *lr = a.evpar + (size_t)(im.evoff + eR[u]) * a.evw;
            same = ls[0] == lr[0];
            for (int j = 1; j <= ls[0] && same; j++) same = ls[j] == lr[j];
          }
        }
        same = same && rows_eq;
      }
      if (same) continue;
    }
    while (eq) {
      const int ta = ia < ns ? (int)etS[ia] : 0x7fffffff;
      const int tb = ib < nr ? (int)etR[ib] : 0x7fffffff;
      const int td = id < nd ? drows[id] : 0x7fffffff;
      const int t = min(ta, min(tb, td));
      if (t == 0x7fffffff) break;
      ia += ta == t;
      ib += tb == t;
      id += td == t;
      const uint16_t *ls, *lr;
      int cs = row_list(a, im, G, s, t, esS, etS, ns, &ps, &ls);
      int cr = row_list(a, im, G, r, t, esR, etR, nr, &pr, &lr);
      if (cs != cr) eq = false;
      for (int i = 0; i < cs && eq; i++) eq = ls[i] == lr[i];
    }
    if (!eq) atomicOr(&flag, 2);
  }
  __syncthreads();
  if (flag) {
    if (tid == 0) {
      atomicOr(&a.status[item], ST_COLLISION);
      a.js[im.set * a.perms + im.perm] = 0.0;
    }
    return;
  }
  phase_mark(a, 1, tph);
  // unique networks in network order: ulist[u] = s
  for (int s = tid; s < S; s += 256) work[s] = rep[s] == s ? 1 : 0;
  __syncthreads();
  {
    int nu = block_exscan(work, S, wsum);
    for (int s = tid; s < S; s += 256)
      if (rep[s] == s) ulist[work[s]] = s;
    if (tid == 0) {
      U = nu;
      a.uniq_u[item] = nu;
    }
  }
  __syncthreads();
  for (int s = tid; s < S; s += 256) a.rep[im.soff + s] = rep[s];
  const int UU = U;
  phase_mark(a, 2, tph);
  // ---- 3. families of the unique networks (hash table of at least 2x the
  // families an item can hold: in the phase-shared LDS region, which is free
  // between the network table and the BDeu stage, when it fits; else in
  // global memory)
  // keys inserted: the 2G full-network rows and the events of unique
  // networks (EU of them)
  __shared__ int EU;
  if (tid == 0) EU = 0;
  __syncthreads();
  {
    int eu = 0;
    for (int u = tid; u < UU; u += 256) eu += evst[ulist[u] + 1] - evst[ulist[u]];
    if (eu) atomicAdd(&EU, eu);
  }
  __syncthreads();
  int FH = 1024;
  while (FH < 2 * (2 * G + EU) && FH < im.fhsize) FH *= 2;
  unsigned long long *fk = a.famht + im.fhoff;
  int *fv = a.famhv + im.fhoff;
  if ((long long)FH * 12 <= (long long)a.score_shared) {
    fk = (unsigned long long *)dyn;
    fv = (int *)(fk + FH);
  }
  for (int i = tid; i < FH; i += 256) {
    fk[i] = 0ULL;
    fv[i] = 0x7fffffff;
  }
  if (tid == 0) F = 0;
  __syncthreads();
  // each key's table slot is kept (famfull for the full rows, evfam for the
  // events) and turned into its family number once the families are numbered
  int *evfam = a.evfam + im.evoff;
  for (int i = tid; i < 2 * G + E; i += 256) {
    uint64_t key;
    if (i < 2 * G) {
      int g = i / G, t = i - g * G;
      size_t lb = (size_t)(im.goff * 2 + g * G + t);
      key = fam_key(t, a.lcnt[lb], a.lfull + lb * MPW);
    } else {
      const int e = i - 2 * G;
      const uint32_t k = ek[e];
      const int s = (int)(k & 0xffffu);
      if (rep[s] != s) {
        evfam[e] = -1;
        continue;
      }
      const uint16_t *pp = epar + (size_t)e * a.evw;
      key = fam_key((int)(k >> 16), pp[0], pp + 1);
    }
    const int h = ht_insert_min(fk, fv, key, 0, FH);
    if (h < 0) atomicOr(&flag, 4);
    if (i < 2 * G)
      famfull[i] = h;
    else
      evfam[i - 2 * G] = h;
  }
  __syncthreads();
  for (int i = tid; i < FH; i += 256) {
    if (fk[i] != 0ULL) {
      int f = atomicAdd(&F, 1);
      fv[i] = f;
      if (f < im.fcap) a.famkey[im.foff + f] = fk[i];
    }
  }
  __syncthreads();
  if (tid == 0) a.famcount[item] = min(F, im.fcap);
  if (flag || F > im.fcap || 2 * F > FH) {
    if (tid == 0) atomicOr(&a.status[item], ST_FAM_OVERFLOW);
    return;
  }
  for (int i = tid; i < 2 * G; i += 256) {
    const int slot = famfull[i];
    famfull[i] = slot >= 0 ? fv[slot] : -1;
  }
  for (int e = tid; e < E; e += 256) {
    const int slot = evfam[e];
    evfam[e] = slot >= 0 ? fv[slot] : -1;
  }
  __syncthreads();
  phase_mark(a, 3, tph);
  // family scores live in global memory (their count is unbounded by LDS)
  double *fscore = a.famscore + im.foff * 2;
  const int NF = F;
  if (a.counters && tid == 0) {
    cnt_add(a, 5, (unsigned long long)NF);      // families
    cnt_add(a, 6, (unsigned long long)UU);      // unique networks
    cnt_add(a, 7, (unsigned long long)E);       // events
  }
  // ---- 4. BDeu per (family, scoring class): kernel5.cu:4-236 -------------
  // each wave scores its own (family, class) pairs: the family's state masks
  // staged in LDS, one histogram cell per lane, the inner sums of the
  // reference's double loop per lane (one lane per parent configuration j),
  // the outer sum over j by lane 0, both in the reference's order
  unsigned long long st_hist = 0, st_bins = 0;                    // algorithmic bytes
  // (a) families with at most one parent (at most 9 cells): one lane each,
  // the same j-outer / k-inner order
  for (int fg = tid; fg < 2 * NF; fg += 256) {
    const int f = fg >> 1, g = fg & 1;
    const uint64_t key = a.famkey[im.foff + f] - 1;
    const int t = (int)(key & 1023), np = (int)((key >> 10) & 3);
    if (np > 1) continue;
    const int p0 = (int)((key >> 12) & 1023);
    const uint8_t *sf = a.sflags + im.goff * 2 + g * G;
    const uint32_t f0 = sf[t], f1 = np > 0 ? sf[p0] : 1u;
    const int r = __popc(f0), q = __popc(f1);
    double sum2 = 0.0;
    if (r > 0 && q > 0) {
      const int W = a.bitw;
      const uint64_t *mk = a.bits + (size_t)(im.goff * 2 + g * G) * 3 * W;
      const double *lq = a.lgtab + (size_t)q * a.lgstride;
      const double *lqr = a.lgtab + (size_t)(q * r) * a.lgstride;
      const double lq0 = lq[0], lqr0 = lqr[0];
      if (np == 1 && !a.plain) {
        // one parent: the (child, parent) histogram in class g is the pair's
        // class-g tally, cell 3 x + y for state x of the pair's higher gene
        // and y of its lower one (the cutoff schedule stores class 1's)
        uint64_t w0, w1, w2;
        class_tally_words(a, im, sm, t > p0 ? pair_index(t, p0) : pair_index(p0, t), g, w0, w1, w2);
        for (int j = 0; j < q; j++) {
          const int sj = nth_state(f1, j);
          int nij = 0;
          double sum1 = 0.0;
          for (int k = 0; k < r; k++) {
            const int sk = nth_state(f0, k);
            const int cell = t > p0 ? sk * 3 + sj : sj * 3 + sk;
            const uint64_t wv = cell < 4 ? w0 : cell < 8 ? w1 : w2;
            const int cnt = (int)((wv >> (16 * (cell & 3))) & 0xffffu);
            nij += cnt;
            sum1 = sum1 + (lqr[cnt] - lqr0);
          }
          sum2 = sum2 + ((lq0 - lq[nij]) + sum1);
        }
      } else {
        for (int j = 0; j < q; j++) {
          const uint64_t *mp_ = np > 0 ? mk + (size_t)(p0 * 3 + nth_state(f1, j)) * W : nullptr;
          int nij = 0;
          double sum1 = 0.0;
          for (int k = 0; k < r; k++) {
            const uint64_t *mc = mk + (size_t)(t * 3 + nth_state(f0, k)) * W;
            int cnt = 0;
            for (int w = 0; w < W; w++) cnt += __popcll(np > 0 ? (mc[w] & mp_[w]) : mc[w]);
            nij += cnt;
            sum1 = sum1 + (lqr[cnt] - lqr0);
          }
          sum2 = sum2 + ((lq0 - lq[nij]) + sum1);
        }
      }
      st_bins += (unsigned long long)(4 * q * r);
    }
    fscore[f * 2 + g] = sum2;
    st_hist += (unsigned long long)(1 + np) * (unsigned long long)(g ? a.n2 : a.n1);
  }
  // (b) families with two or three parents: appended to the launch's task
  // list for the BDeu kernel (its lane groups over all items of the chunk),
  // or, when the list is full, scored here by this block's lane groups
  {
    __shared__ int nb, bbase, bpos;
    if (tid == 0) nb = bpos = 0;
    __syncthreads();
    int mine_n = 0;
    for (int fg = tid; fg < 2 * NF; fg += 256) {
      const uint64_t key = a.famkey[im.foff + (fg >> 1)] - 1;
      mine_n += ((key >> 10) & 3) > 1;
    }
    if (mine_n) atomicAdd(&nb, mine_n);
    __syncthreads();
    if (tid == 0) {
      bbase = -1;
      if (a.bd_tasks && nb > 0) {
        const unsigned b0 = atomicAdd(a.bd_count, (unsigned)nb);
        if ((long long)b0 + nb <= (long long)a.bd_cap) bbase = (int)b0;
      }
    }
    __syncthreads();
    if (bbase >= 0) {
      for (int fg = tid; fg < 2 * NF; fg += 256) {
        const uint64_t key = a.famkey[im.foff + (fg >> 1)] - 1;
        if (((key >> 10) & 3) > 1) a.bd_tasks[bbase + atomicAdd(&bpos, 1)] = make_int2(item, fg);
      }
    } else if (nb > 0) {
      const int G4 = a.score_groups, GS = 64 / G4, gi = lane / GS, gl = lane - gi * GS;
      uint64_t *gmst;
      uint16_t *ghist;
      double *glgv;
      bdeu_stage(dyn, G4, a.bitw, wave, gi, &gmst, &ghist, &glgv);
      for (int fg0 = wave * G4; fg0 < 2 * NF; fg0 += EDDY_NWAVES * G4) {
        const int fg = fg0 + gi;
        bdeu_group(a, im, G, fg < 2 * NF ? (fg >> 1) : -1, fg & 1, gl, GS, gmst, ghist, glgv, st_hist, st_bins);
      }
    }
  }
  if (a.counters) {
    __shared__ unsigned long long sh_st[2];
    if (tid == 0) sh_st[0] = sh_st[1] = 0ULL;
    __syncthreads();
    if (st_hist) atomicAdd(&sh_st[0], st_hist);
    if (st_bins) atomicAdd(&sh_st[1], st_bins);
    __syncthreads();
    if (tid == 0) {
      cnt_add(a, 4, sh_st[0]);  // histogram state bytes read (reference algorithm)
      cnt_add(a, 8, sh_st[1]);  // histogram bin bytes
    }
  }
  // what the tail kernel needs besides the family scores
  for (int u = tid; u < UU; u += 256) a.sc_ulist[im.soff + u] = ulist[u];
  for (int x = tid; x < S; x += 256) a.sc_evst[im.soff + x] = evst[x];
  for (int i = tid; i < 2 * G; i += 256) a.sc_famfull[im.goff * 2 + i] = famfull[i];
  __syncthreads();
  phase_mark(a, 4, tph);
}

__global__ __launch_bounds__(256) void k_score(KArgs a) { score_body<false>(a); }
__global__ __launch_bounds__(256) void k_score_g(KArgs a) { score_body<true>(a); }

// Network scores, likelihoods and JS of every item whose score kernel ran to
// the end, once every BDeu task of the launch has finished.  One workgroup per
// item; the unique-network list, event ranges and full families come from
// the score kernel through global memory.
// GM: the network scores / likelihoods in the item's likelihood slots
// (a.lval) and the KL terms in its global scratch instead of LDS
template <bool GM>
__device__ __forceinline__ void score_tail_body(const KArgs &a) {
  const int item = a.order[blockIdx.x];
  if (a.status[item] & (ST_FAM_OVERFLOW | ST_COLLISION)) return;
  const ItemMeta im = a.items[item];
  const SetMeta sm = a.sets[im.set];
  const int G = sm.G, S = a.S, n1 = a.n1;
  const int tid = threadIdx.x;
  extern __shared__ __align__(16) unsigned char dyn[];
  const int UU = a.uniq_u[item];
  const int E = min(a.evcount[item], im.evcap);
  const int *es = a.evorder + im.evoff;
  const uint16_t *et = a.evt + im.evoff;  // rows of es[], sorted per network
  const int *ulist = a.sc_ulist + im.soff;
  const int *evst = a.sc_evst + im.soff;
  const int *famfull = a.sc_famfull + im.goff * 2;
  const double *fscore = a.famscore + im.foff * 2;
  unsigned long long tph = 0;
  phase_mark(a, -1, tph);
  // ---- 5. network scores, in node order (main.cu:1335-1342) -------------
  // base[(gs * 2 + g) * G + t]: the score in class g of row t's family in
  // the full network of class gs (every network's value outside its event
  // rows), staged in LDS; event rows read their own family's score
  double *base = (double *)dyn;
  for (int i = tid; i < 4 * G; i += 256) {
    const int q = i / G, t = i - q * G, gs = q >> 1, g = q & 1;
    const int f = famfull[gs * G + t];
    base[i] = f >= 0 ? fscore[f * 2 + g] : eddy_nan();
  }
  __syncthreads();
  // network scores, then (in place) the likelihoods, in LDS after the
  // phase-shared region: [g*S + u]
  double *dist = GM ? a.lval + im.soff * 2 : (double *)(dyn + a.tail_shared);
  for (int ug = tid; ug < 2 * UU; ug += 256) {
    const int u = ug >> 1, g = ug & 1;
    const int s = ulist[u];
    const int b = evst[s], ne = (s + 1 < S ? evst[s + 1] : E) - b;
    const double *bs = base + (net_class(s, n1) * 2 + g) * G;
    int ep = 0, te = ne > 0 ? (int)et[b] : G;
    double d = 0.0;
    for (int t = 0; t < G; t++) {
      double v = bs[t];
      if (t == te) {
        const int f = a.evfam[im.evoff + es[b + ep]];
        v = f >= 0 ? fscore[f * 2 + g] : eddy_nan();
        ep++;
        te = ep < ne ? (int)et[b + ep] : G;
      }
      d = d + v;
    }
    dist[g * S + u] = d;
  }
  __syncthreads();
  phase_mark(a, 5, tph);
  // ---- 6. likelihoods (main.cu:1294-1404) and JS (routines.cu:42-60) ----
  __shared__ double sh_inalpha[2], sh_scale[2], sh_nonInf[2], sh_infc[2];
  if (tid < 2) {
    int g = tid;
    double mn = dist[g * S], mx = dist[g * S], sum = 0.0;
    for (int u = 0; u < UU; u++) {
      double d = dist[g * S + u];
      sum = sum + d;
      if (d > mx) mx = d;
      if (d < mn) mn = d;
    }
    sh_inalpha[g] = -1 * (sum / UU);
    sh_scale[g] = 10 / (mx - mn);
  }
  __syncthreads();
  double *lk = dist;  // in place: dist -> exp(adjusted)
  for (int ug = tid; ug < 2 * UU; ug += 256) {
    int g = ug / UU, u = ug - g * UU;
    lk[g * S + u] = eddy_exp((dist[g * S + u] + sh_inalpha[g]) * sh_scale[g]);
  }
  __syncthreads();
  if (tid < 2) {
    int g = tid;
    double nonInf = 0.0, infc = 0.0;
    for (int u = 0; u < UU; u++) {
      double l = lk[g * S + u];
      nonInf = nonInf + l;
      if (l >= eddy_inf() || l <= -eddy_inf()) infc++;
    }
    sh_nonInf[g] = nonInf;
    sh_infc[g] = infc;
  }
  __syncthreads();
  for (int ug = tid; ug < 2 * UU; ug += 256) {
    int g = ug / UU, u = ug - g * UU;
    double l = lk[g * S + u];
    double isinf = (l >= eddy_inf() || l <= -eddy_inf()) ? 1.0 : 0.0;
    lk[g * S + u] = sh_infc[g] == 0 ? l / sh_nonInf[g] : isinf / sh_infc[g];
  }
  __syncthreads();
  // KL terms in parallel (the family-score space is free now); summed
  // serially in network order.  A skipped term (M = 0 or P = 0,
  // routines.cu:47-53) adds +0.0, which leaves every partial sum unchanged:
  // the sums start at +0.0 and no term is -0.0.
  double *kt = GM ? (double *)(a.gscr + (size_t)item * a.gscr_item) : (double *)dyn;
  for (int ug = tid; ug < 2 * UU; ug += 256) {
    int g = ug / UU, u = ug - g * UU;
    double p1 = lk[u], p2 = lk[S + u];
    double m = (p1 + p2) / 2;
    double pp = g ? p2 : p1;
    kt[ug] = (m == 0 || pp == 0) ? 0.0 : pp * eddy_log(pp / m);
  }
  // the likelihoods for the permutation-0 captures (with GM they are there)
  if (!GM) {
    double *lv = a.lval + im.soff * 2;
    for (int ug = tid; ug < 2 * UU; ug += 256) {
      const int g = ug / UU, u = ug - g * UU;
      lv[g * S + u] = lk[g * S + u];
    }
  }
  __syncthreads();
  if (tid == 0) {
    double k1 = 0.0, k2 = 0.0;
    for (int u = 0; u < UU; u++) k1 = k1 + kt[u];
    for (int u = 0; u < UU; u++) k2 = k2 + kt[UU + u];
    double js = k1 / 2 + k2 / 2;
    a.js[im.set * a.perms + im.perm] = js / eddy_log(2.0);
  }
  phase_mark(a, 6, tph);
}

__global__ __launch_bounds__(256) void k_score_tail(KArgs a) { score_tail_body<false>(a); }
__global__ __launch_bounds__(256) void k_score_tail_g(KArgs a) { score_tail_body<true>(a); }

// BDeu of every (family, class) task the launch's score kernels listed, in
// lane groups (G4 per wave, as in the score kernel), the blocks striding over
// the list.  Each task writes its own family score.
__global__ __launch_bounds__(256) void k_bdeu(KArgs a) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int G4 = a.score_groups, GS = 64 / G4, gi = lane / GS, gl = lane - gi * GS;
  extern __shared__ __align__(16) unsigned char dyn[];
  uint64_t *gmst;
  uint16_t *ghist;
  double *glgv;
  bdeu_stage(dyn, G4, a.bitw, wave, gi, &gmst, &ghist, &glgv);
  const long long ntask = min((long long)*a.bd_count, (long long)a.bd_cap);
  unsigned long long st_hist = 0, st_bins = 0;
  const long long per_wave = G4, stride = (long long)gridDim.x * EDDY_NWAVES * G4;
  for (long long t0 = ((long long)blockIdx.x * EDDY_NWAVES + wave) * per_wave; t0 < ntask; t0 += stride) {
    const long long tk = t0 + gi;
    int item = 0, fg = -2;
    if (tk < ntask) {
      const int2 w = a.bd_tasks[tk];
      item = w.x;
      fg = w.y;
    }
    const ItemMeta im = a.items[item];
    const int G = a.sets[im.set].G;
    bdeu_group(a, im, G, fg >= 0 ? (fg >> 1) : -1, fg & 1, gl, GS, gmst, ghist, glgv, st_hist, st_bins);
  }
  if (a.counters) {
    __shared__ unsigned long long sh_st[2];
    if (tid == 0) sh_st[0] = sh_st[1] = 0ULL;
    __syncthreads();
    if (st_hist) atomicAdd(&sh_st[0], st_hist);
    if (st_bins) atomicAdd(&sh_st[1], st_bins);
    __syncthreads();
    if (tid == 0) {
      cnt_add(a, 4, sh_st[0]);
      cnt_add(a, 8, sh_st[1]);
    }
  }
}

// --------------------------------------------------------------------------
// K5: p-value per set: routines.cu:494-508, main.cu:1465-1502
// out[set*6 + ...] = {status, js0, pDDN, alpha, beta, perms_run}
// --------------------------------------------------------------------------
__global__ void k_pvalue(const double *js, int nsets, int perms, const int *nan0, double *out) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsets) return;
  const double *v = js + (size_t)s * perms;
  double *o = out + (size_t)s * 6;
  if (nan0[s]) {
    o[0] = 1.0;
    o[1] = -999.0;
    o[2] = eddy_nan();
    o[3] = o[4] = eddy_nan();
    o[5] = 1;
    return;
  }
  int nan = 0;
  for (int i = 0; i < perms; i++) nan |= eddy_isnan(v[i]);
  o[1] = v[0];
  o[5] = perms;
  if (v[0] < 0 || nan) {
    o[0] = 1.0;
    o[2] = o[3] = o[4] = eddy_nan();
    return;
  }
  double sum = 0.0;
  for (int i = 0; i < perms; i++) sum = sum + v[i];
  double mu = sum / (double)perms;
  double vs = 0.0;
  for (int i = 0; i < perms; i++) {
    double d = v[i] - mu;
    vs = vs + d * d;
  }
  double var = vs / (double)perms;
  double al = (((1 - mu) / var) - (1 / mu)) * (mu * mu);
  double be = al * (1 / mu - 1);
  int de = 0;
  double p = eddy_ibetac(al, be, v[0], &de);
  o[0] = de ? 2.0 : 0.0;
  o[2] = p;
  o[3] = al;
  o[4] = be;
}

// gather the rows of every set's genes (row stride NP, padded with state 3)
__global__ void k_gather(const uint8_t *x, int N, int NP, const int *genes, const long long *rowoff, int nrows,
                         uint8_t *xs) {
  const int r = blockIdx.x;
  if (r >= nrows) return;
  const uint4 *src = (const uint4 *)(x + (size_t)genes[r] * NP);
  uint4 *dst = (uint4 *)(xs + rowoff[r]);
  for (int i = threadIdx.x; i < NP / 16; i += blockDim.x) dst[i] = src[i];
}

// permutation-0 captures: one workgroup per slice at a time, 4-byte words
// when source, destination and length allow it, else bytes
__global__ __launch_bounds__(256) void k_copy_slices(const Slice *sl, int n) {
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const Slice q = sl[i];
    const uintptr_t al = (uintptr_t)q.src | (uintptr_t)q.dst | (uintptr_t)q.bytes;
    if ((al & 3) == 0) {
      const uint32_t *s = (const uint32_t *)q.src;
      uint32_t *d = (uint32_t *)q.dst;
      for (long long k = threadIdx.x; k < q.bytes / 4; k += blockDim.x) d[k] = s[k];
    } else {
      const uint8_t *s = (const uint8_t *)q.src;
      uint8_t *d = (uint8_t *)q.dst;
      for (long long k = threadIdx.x; k < q.bytes; k += blockDim.x) d[k] = s[k];
    }
  }
}

// permutation-0 JS of sets first..first+n-1 (one value per row of js)
__global__ void k_js0(const double *js, int perms, int first, int n, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = js[(size_t)(first + i) * perms];
}

// Learned window per (set, row): the first column at which the expected number
// of full-test edges (the network's scores, summed in column order) reaches
// mp + kappa; the whole row when it never does.  One block per set.
__global__ void k_window(const SetMeta *sets, int nsets, const float *score, int mp, float kappa, int *win) {
  const int s = blockIdx.x;
  if (s >= nsets) return;
  const SetMeta sm = sets[s];
  for (int t = threadIdx.x; t < sm.G; t += blockDim.x) {
    float acc = 0.f;
    int W = t - 1;
    const float *sc = score + sm.pairbase + (long long)t * (t - 1) / 2;
    for (int col = 0; col < t; col++) {
      acc += sc[col];
      if (acc >= (float)mp + kappa) {
        W = col;
        break;
      }
    }
    win[sm.genebase + t] = W;
  }
}
