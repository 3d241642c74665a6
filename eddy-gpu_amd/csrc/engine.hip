// engine.hip - host side of libeddy: context, dataset upload, batching of
// (gene set, permutation) items, the two-phase permutation stream, launches
// of the gfx950 kernels and the C ABI declared in include/eddy.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "eddy.h"
#include "engine_types.h"
#include "perm_stream.h"

// kernels (kernels.hip)
__global__ void k_count_test(KArgs a);
__global__ void k_loo(KArgs a);
__global__ void k_rows(KArgs a);
__global__ void k_score(KArgs a);
__global__ void k_pvalue(const double *js, int nsets, int perms, const int *nan0, double *out);
__global__ void k_gather(const uint8_t *x, int N, int NP, const int *genes, const long long *rowoff, int nrows,
                         uint8_t *xs);

namespace {

struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
};

// Per-set copy of what permutation 0 leaves behind (the reference's first_*
// arrays and edgeListData1/2, main.cu:782-819, 1085-1195, 1406-1409).
struct Perm0 {
  int U = 0;
  std::vector<uint8_t> lcnt;    // [2G]
  std::vector<uint16_t> lfull;  // [2G * EDDY_MAX_MP]
  std::vector<int> rep;         // [S]
  std::vector<Event> ev;        // events of the permutation
  std::vector<double> lval;     // [2S]: g*S + u
  std::vector<uint32_t> info;   // [c]
  std::vector<uint64_t> famkey; // distinct families (t, first <= 3 parents)
  std::vector<double> famscore; // [f*2 + g] BDeu of family f on class g
};

}  // namespace

struct eddy_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // dataset
  int genes = 0, N = 0, n1 = 0, n2 = 0;
  uint8_t *dX = nullptr;
  std::vector<uint8_t> classOf;
  double *dLG = nullptr;
  int lgstride = 0;
  double lg15 = 0, lg2 = 0, lg3 = 0;
  PermStream stream_gen;
  // device buffers, grown on demand
  DevBuf b_sets, b_xs, b_prior, b_items, b_labels, b_tiles, b_info, b_tab, b_lfull, b_lcnt, b_sflags, b_ev,
      b_evcount, b_hnet, b_rep, b_clsg, b_evorder, b_evfam, b_famkey, b_famscore, b_lval, b_uniq, b_js,
      b_status, b_tests, b_genes, b_rowoff, b_nan0, b_pout, b_famcount, b_counters;
  // last run
  int last_perms = 0;
  std::vector<int> last_G;
  std::vector<std::vector<double>> last_js;
  std::vector<Perm0> last_p0;
  eddy_timing timing{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

#define HIPCHK(ctx, x)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) {                                                             \
      (ctx)->err = std::string(#x) + ": " + hipGetErrorString(e_);                      \
      return EDDY_ERR_HIP;                                                              \
    }                                                                                   \
  } while (0)

static int grow(eddy_ctx *ctx, DevBuf &b, size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (b.n >= bytes) return EDDY_OK;
  if (b.p) HIPCHK(ctx, hipFree(b.p));
  b.p = nullptr;
  b.n = 0;
  size_t want = bytes + bytes / 4;
  HIPCHK(ctx, hipMalloc(&b.p, want));
  b.n = want;
  return EDDY_OK;
}

#define GROW(b, bytes)                                 \
  do {                                                 \
    int rc_ = grow(ctx, ctx->b, (size_t)(bytes));      \
    if (rc_) return rc_;                               \
  } while (0)

#define CHK(x)              \
  do {                      \
    int rc_ = (x);          \
    if (rc_) return rc_;    \
  } while (0)

extern "C" const char *eddy_version(void) { return "libeddy 0.1 (gfx950)"; }

extern "C" int eddy_ctx_create(int device, eddy_ctx **out) {
  if (!out) return EDDY_ERR_ARG;
  *out = nullptr;
  if (device == -1) {  // host-only context: permutation stream only
    eddy_ctx *ctx = new eddy_ctx();
    ctx->device = -1;
    *out = ctx;
    return EDDY_OK;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return EDDY_ERR_NODEV;
  if (device < 0 || device >= n) return EDDY_ERR_ARG;
  eddy_ctx *ctx = new eddy_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
    delete ctx;
    return EDDY_ERR_HIP;
  }
  *out = ctx;
  return EDDY_OK;
}

extern "C" int eddy_ctx_destroy(eddy_ctx *ctx) {
  if (!ctx) return EDDY_ERR_ARG;
  if (ctx->device < 0) {
    delete ctx;
    return EDDY_OK;
  }
  hipSetDevice(ctx->device);
  DevBuf *all[] = {&ctx->b_sets, &ctx->b_xs, &ctx->b_prior, &ctx->b_items, &ctx->b_labels, &ctx->b_tiles,
                   &ctx->b_info, &ctx->b_tab, &ctx->b_lfull, &ctx->b_lcnt, &ctx->b_sflags, &ctx->b_ev,
                   &ctx->b_evcount, &ctx->b_hnet, &ctx->b_rep, &ctx->b_clsg, &ctx->b_evorder, &ctx->b_evfam,
                   &ctx->b_famkey, &ctx->b_famscore, &ctx->b_lval, &ctx->b_uniq, &ctx->b_js, &ctx->b_status,
                   &ctx->b_tests, &ctx->b_genes, &ctx->b_rowoff, &ctx->b_nan0, &ctx->b_pout, &ctx->b_famcount, &ctx->b_counters};
  for (DevBuf *b : all)
    if (b->p) hipFree(b->p);
  if (ctx->dX) hipFree(ctx->dX);
  if (ctx->dLG) hipFree(ctx->dLG);
  if (ctx->ev0) hipEventDestroy(ctx->ev0);
  if (ctx->ev1) hipEventDestroy(ctx->ev1);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return EDDY_OK;
}

extern "C" const char *eddy_last_error(const eddy_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

extern "C" int eddy_stream_reset(eddy_ctx *ctx) {
  if (!ctx) return EDDY_ERR_ARG;
  ctx->stream_gen.reset(1);
  return EDDY_OK;
}

extern "C" int eddy_stream_draw(eddy_ctx *ctx, int n, int32_t *out) {
  if (!ctx || n < 0 || (!out && n)) return EDDY_ERR_ARG;
  for (int i = 0; i < n; i++) out[i] = (int32_t)ctx->stream_gen.next();
  return EDDY_OK;
}

extern "C" int eddy_stream_permutation(eddy_ctx *ctx, int samples, int n1, int perm, uint8_t *labels) {
  if (!ctx || samples < 1 || n1 < 0 || n1 > samples || !labels) return EDDY_ERR_ARG;
  std::vector<int> ord;
  ctx->stream_gen.permutation(samples, ord);
  if (perm > 0)
    for (int i = 0; i < samples; i++) labels[ord[i]] = i < n1 ? 0 : 1;  // main.cu:561-570
  return EDDY_OK;
}

extern "C" int eddy_upload_dataset(eddy_ctx *ctx, const uint8_t *x, int genes, int samples, const uint8_t *class_of) {
  if (!ctx || !x || !class_of || genes <= 0 || samples <= 1) return EDDY_ERR_ARG;
  if (ctx->device < 0) return EDDY_ERR_NODEV;
  if (samples > EDDY_MAX_N) {
    ctx->err = "more than 2046 samples";
    return EDDY_ERR_ARG;
  }
  hipSetDevice(ctx->device);
  int n1 = 0;
  for (int s = 0; s < samples; s++) {
    if (class_of[s] > 1) return EDDY_ERR_ARG;
    n1 += class_of[s] == 0;
  }
  if (n1 == 0 || n1 == samples) {
    ctx->err = "both classes need samples";
    return EDDY_ERR_ARG;
  }
  const int NP = (samples + 63) & ~63;
  std::vector<uint8_t> padded((size_t)genes * NP, 3);
  for (int g = 0; g < genes; g++)
    for (int s = 0; s < samples; s++) {
      uint8_t v = x[(size_t)g * samples + s];
      padded[(size_t)g * NP + s] = v > 2 ? 3 : v;
    }
  if (ctx->dX) HIPCHK(ctx, hipFree(ctx->dX));
  ctx->dX = nullptr;
  HIPCHK(ctx, hipMalloc(&ctx->dX, padded.size()));
  HIPCHK(ctx, hipMemcpy(ctx->dX, padded.data(), padded.size(), hipMemcpyHostToDevice));
  ctx->genes = genes;
  ctx->N = samples;
  ctx->n1 = n1;
  ctx->n2 = samples - n1;
  ctx->classOf.assign(class_of, class_of + samples);
  // lgamma(n + 4/d) for d = 1..81, n = 0..N: the BDeu arguments of
  // kernel5.cu:205-216 (nijk + N'/(q r), nij + N'/q with N' = 4)
  ctx->lgstride = samples + 1;
  std::vector<double> lg((size_t)82 * ctx->lgstride, 0.0);
  for (int d = 1; d <= 81; d++)
    for (int n = 0; n <= samples; n++) lg[(size_t)d * ctx->lgstride + n] = std::lgamma((double)n + 4.0 / (double)d);
  if (ctx->dLG) HIPCHK(ctx, hipFree(ctx->dLG));
  ctx->dLG = nullptr;
  HIPCHK(ctx, hipMalloc(&ctx->dLG, lg.size() * sizeof(double)));
  HIPCHK(ctx, hipMemcpy(ctx->dLG, lg.data(), lg.size() * sizeof(double), hipMemcpyHostToDevice));
  ctx->lg15 = std::lgamma(1.5);
  ctx->lg2 = std::lgamma(2.0);
  ctx->lg3 = std::lgamma(3.0);
  return EDDY_OK;
}

namespace {

// ---- certainty intervals of the edge decision (see edgetest.h) ------------
// upper tail 1 - P(p, x) for p = 1/2, 1, 2 (closed forms)
long double q_tail(int pi, long double x) {
  if (pi == 0) return erfcl(sqrtl(x));
  if (pi == 1) return expl(-x);
  return expl(-x) * (1.0L + x);
}

// boundary x where the decreasing tail crosses target: returns (lo, hi)
// bracketing it to long-double precision
void tail_root(int pi, long double target, long double &lo, long double &hi) {
  lo = 0.0L;
  hi = 4000.0L;
  for (int i = 0; i < 400; i++) {
    long double mid = 0.5L * (lo + hi);
    if (q_tail(pi, mid) > target) lo = mid; else hi = mid;
  }
}

const double kPs[3] = {0.5, 1.0, 2.0};

void edge_params(eddy_ctx *ctx, const eddy_params *p, eddy_edge_params &ep) {
  ep.aE = p->alpha_edge;
  ep.aEP = p->alpha_edge_prior;
  ep.lg15 = ctx->lg15;
  ep.lg2 = ctx->lg2;
  ep.lg3 = ctx->lg3;
  ep.p_adjust = p->p_adjust;
  // shortcuts only for thresholds strictly inside (0, 1)
  ep.shortcut = (p->alpha_edge > 1e-300 && p->alpha_edge < 1.0 && p->alpha_edge_prior < 1.0) ? 1 : 0;
  const long double L = (long double)eddy_log(1.0e-37);
  const double lg[3] = {ctx->lg15, ctx->lg2, ctx->lg3};
  for (int pi = 0; pi < 3; pi++) {
    // AS 147 underflow cut: p ln x - lgamma(p+1) - x = ln 1e-37, decreasing for x > p
    long double lo = kPs[pi] + 1.0L, hi = 4000.0L;
    for (int i = 0; i < 400; i++) {
      long double mid = 0.5L * (lo + hi);
      long double g = (long double)kPs[pi] * logl(mid) - (long double)lg[pi] - mid - L;
      if (g > 0) lo = mid; else hi = mid;
    }
    ep.xclo[pi] = (double)(lo - 1e-6L);
    ep.xchi[pi] = (double)(hi + 1e-6L);
  }
}

void edge_bounds(const eddy_params *p, int c, eddy_edge_bounds &eb) {
  const double m = 1e-5;
  for (int pr = 0; pr < 2; pr++) {
    double alpha = pr ? std::max(p->alpha_edge, p->alpha_edge_prior) : p->alpha_edge;
    double thr = p->p_adjust ? alpha / (double)std::max(c, 1) : alpha;
    for (int pi = 0; pi < 3; pi++) {
      long double lo, hi;
      eb.xlo[pr][pi] = -1.0;
      eb.xhi[pr][pi] = 1e300;
      if (thr > 1e-12 && thr * (1 + m) < 1.0) {
        tail_root(pi, (long double)thr * (1.0L + m), lo, hi);
        eb.xlo[pr][pi] = (double)(lo * (1.0L - 1e-9L));
      }
      if (thr - 5e-9 > 1e-12) {
        tail_root(pi, ((long double)thr - 5e-9L) * (1.0L - m), lo, hi);
        eb.xhi[pr][pi] = (double)(hi * (1.0L + 1e-9L));
      }
    }
  }
}

struct SetPrep {
  std::vector<SetMeta> meta;
  std::vector<uint32_t> prior;
};

int check_params(eddy_ctx *ctx, const eddy_params *p) {
  if (!p || p->mp < 1 || p->mp > EDDY_MAX_MP || p->perms < 1) {
    ctx->err = "invalid parameters (mp must be 1..16, perms >= 1)";
    return EDDY_ERR_ARG;
  }
  return EDDY_OK;
}

// one chunk of items through K1..K4
struct Chunk {
  std::vector<ItemMeta> items;
  std::vector<uint32_t> labels;
  std::vector<uint2> tiles;
  long long pairs = 0, gslots = 0, sslots = 0;
};

int run_chunk(eddy_ctx *ctx, const std::vector<SetMeta> &sets, Chunk &ch, const eddy_params *p, int perms,
              bool capture, std::vector<Perm0> *p0, std::vector<long long> &tests_out) {
  const int N = ctx->N, S = N + 2;
  const int nitems = (int)ch.items.size();
  if (nitems == 0) return EDDY_OK;
  int ev_cap = std::max(1024, 2 * S);
  int maxG = 0;
  for (const ItemMeta &im : ch.items) maxG = std::max(maxG, sets[im.set].G);
  int ht = 1024;
  while (ht < 2 * std::max(S, 2 * maxG + 128)) ht *= 2;
  for (int attempt = 0; attempt < 6; attempt++) {
    int fam_cap = 2 * maxG + ev_cap;
    // offsets
    long long evoff = 0, foff = 0;
    for (int i = 0; i < nitems; i++) {
      ch.items[i].evoff = evoff;
      ch.items[i].foff = foff;
      evoff += ev_cap;
      foff += fam_cap;
    }
    GROW(b_items, sizeof(ItemMeta) * nitems);
    GROW(b_labels, sizeof(uint32_t) * ch.labels.size());
    GROW(b_tiles, sizeof(uint2) * ch.tiles.size());
    GROW(b_info, sizeof(uint32_t) * ch.pairs);
    const bool plain_tests = p->schedule == EDDY_TESTS_ALL;
    if (!plain_tests) GROW(b_tab, sizeof(uint16_t) * 18 * ch.pairs);
    GROW(b_lfull, sizeof(uint16_t) * EDDY_MAX_MP * 2 * ch.gslots);
    GROW(b_lcnt, 2 * ch.gslots);
    GROW(b_sflags, 2 * ch.gslots);
    GROW(b_ev, sizeof(Event) * evoff);
    GROW(b_evcount, sizeof(int) * nitems);
    GROW(b_hnet, sizeof(uint64_t) * ch.sslots);
    GROW(b_rep, sizeof(int) * ch.sslots);
    GROW(b_clsg, sizeof(uint16_t) * ch.sslots);
    GROW(b_evorder, sizeof(int) * evoff);
    GROW(b_evfam, sizeof(int) * evoff);
    GROW(b_famkey, sizeof(uint64_t) * foff);
    GROW(b_famscore, sizeof(double) * 2 * foff);
    GROW(b_lval, sizeof(double) * 2 * ch.sslots);
    GROW(b_uniq, sizeof(int) * nitems);
    GROW(b_famcount, sizeof(int) * nitems);
    GROW(b_status, sizeof(int) * nitems);
    GROW(b_tests, sizeof(long long) * nitems);
    hipStream_t st = ctx->stream;
    HIPCHK(ctx, hipMemcpyAsync(ctx->b_items.p, ch.items.data(), sizeof(ItemMeta) * nitems, hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(ctx->b_labels.p, ch.labels.data(), sizeof(uint32_t) * ch.labels.size(),
                               hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(ctx->b_tiles.p, ch.tiles.data(), sizeof(uint2) * ch.tiles.size(), hipMemcpyHostToDevice,
                               st));
    HIPCHK(ctx, hipMemsetAsync(ctx->b_evcount.p, 0, sizeof(int) * nitems, st));
    HIPCHK(ctx, hipMemsetAsync(ctx->b_status.p, 0, sizeof(int) * nitems, st));
    HIPCHK(ctx, hipMemsetAsync(ctx->b_tests.p, 0, sizeof(long long) * nitems, st));

    KArgs a{};
    a.N = N;
    a.n1 = ctx->n1;
    a.n2 = ctx->n2;
    a.S = S;
    a.mp = p->mp;
    a.p_adjust = p->p_adjust;
    a.plain = plain_tests ? 1 : 0;
    a.ev_cap = ev_cap;
    a.fam_cap = fam_cap;
    a.ht_size = ht;
    a.plan_extra = nullptr;
    a.counters = (unsigned long long *)ctx->b_counters.p;
    edge_params(ctx, p, a.ep);
    if (getenv("EDDY_NO_SHORTCUT")) a.ep.shortcut = 0;
    a.lgtab = ctx->dLG;
    a.lgstride = ctx->lgstride;
    a.hseed = 0x243F6A8885A308D3ULL + (uint64_t)attempt * 0x9E3779B97F4A7C15ULL;
    a.sets = (const SetMeta *)ctx->b_sets.p;
    a.xs = (const uint8_t *)ctx->b_xs.p;
    a.prior = (const uint32_t *)ctx->b_prior.p;
    a.items = (const ItemMeta *)ctx->b_items.p;
    a.labels = (const uint32_t *)ctx->b_labels.p;
    a.tiles = (const uint32_t *)ctx->b_tiles.p;
    a.info = (uint32_t *)ctx->b_info.p;
    a.tab = (uint16_t *)ctx->b_tab.p;
    a.lfull = (uint16_t *)ctx->b_lfull.p;
    a.lcnt = (uint8_t *)ctx->b_lcnt.p;
    a.sflags = (uint8_t *)ctx->b_sflags.p;
    a.ev = (Event *)ctx->b_ev.p;
    a.evcount = (int *)ctx->b_evcount.p;
    a.hnet = (uint64_t *)ctx->b_hnet.p;
    a.rep = (int *)ctx->b_rep.p;
    a.clsg = (uint16_t *)ctx->b_clsg.p;
    a.evorder = (int *)ctx->b_evorder.p;
    a.evfam = (int *)ctx->b_evfam.p;
    a.famkey = (uint64_t *)ctx->b_famkey.p;
    a.famscore = (double *)ctx->b_famscore.p;
    a.lval = (double *)ctx->b_lval.p;
    a.uniq_u = (int *)ctx->b_uniq.p;
    a.famcount = (int *)ctx->b_famcount.p;
    a.js = (double *)ctx->b_js.p;
    a.perms = perms;
    a.status = (int *)ctx->b_status.p;
    a.tests = (long long *)ctx->b_tests.p;

    hipEvent_t ev[5];
    for (int k = 0; k < 5; k++) hipEventCreate(&ev[k]);
    hipEventRecord(ev[0], st);
    if (!ch.tiles.empty()) k_count_test<<<(unsigned)ch.tiles.size(), 256, 0, st>>>(a);
    hipEventRecord(ev[1], st);
    if (!plain_tests) k_loo<<<nitems, 256, 0, st>>>(a);
    hipEventRecord(ev[2], st);
    k_rows<<<nitems, 256, 0, st>>>(a);
    hipEventRecord(ev[3], st);
    const size_t score_lds = (size_t)ht * 12 + (size_t)(4 * S + 1 + 2 * maxG) * 4;
    k_score<<<nitems, 256, score_lds, st>>>(a);
    hipEventRecord(ev[4], st);
    HIPCHK(ctx, hipGetLastError());
    std::vector<int> status(nitems);
    HIPCHK(ctx, hipMemcpyAsync(status.data(), ctx->b_status.p, sizeof(int) * nitems, hipMemcpyDeviceToHost, st));
    std::vector<long long> tests(nitems);
    HIPCHK(ctx, hipMemcpyAsync(tests.data(), ctx->b_tests.p, sizeof(long long) * nitems, hipMemcpyDeviceToHost, st));
    HIPCHK(ctx, hipStreamSynchronize(st));
    const int kidx[4] = {0, 1, 2, 3};
    for (int k = 0; k < 4; k++) {
      if (k == 1 && plain_tests) continue;
      if (k == 0 && ch.tiles.empty()) continue;
      float ms = 0;
      hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
      ctx->timing.kernel_ms[kidx[k]] += ms;
      ctx->timing.launches[kidx[k]] += 1;
    }
    for (int k = 0; k < 5; k++) hipEventDestroy(ev[k]);
    int bad = 0;
    for (int s : status) bad |= s;
    if (bad & ST_NOT_EVALUATED) {
      ctx->err = "internal: a row read a leave-one-out test that was not evaluated";
      return EDDY_ERR_STATE;
    }
    if (bad & (ST_EV_OVERFLOW | ST_FAM_OVERFLOW | ST_COLLISION)) {
      if (bad & ST_EV_OVERFLOW) ev_cap *= 4;
      if (bad & ST_FAM_OVERFLOW) ht *= 2;
      continue;  // retry the chunk (larger capacity / new hash seed)
    }
    for (int i = 0; i < nitems; i++) {
      tests_out[i] = tests[i];
      ctx->timing.tests += tests[i];
    }
    ctx->timing.items += nitems;
    ctx->timing.pairs += ch.pairs;
    if (capture) {
      // copy back permutation-0 state of every item (one item per set)
      std::vector<int> evc(nitems), uq(nitems), fc(nitems);
      HIPCHK(ctx, hipMemcpy(evc.data(), ctx->b_evcount.p, sizeof(int) * nitems, hipMemcpyDeviceToHost));
      HIPCHK(ctx, hipMemcpy(fc.data(), ctx->b_famcount.p, sizeof(int) * nitems, hipMemcpyDeviceToHost));
      HIPCHK(ctx, hipMemcpy(uq.data(), ctx->b_uniq.p, sizeof(int) * nitems, hipMemcpyDeviceToHost));
      for (int i = 0; i < nitems; i++) {
        const ItemMeta &im = ch.items[i];
        const SetMeta &sm = sets[im.set];
        Perm0 &q = (*p0)[im.set];
        q.U = uq[i];
        q.lcnt.resize(2 * sm.G);
        q.lfull.resize((size_t)2 * sm.G * EDDY_MAX_MP);
        q.rep.resize(S);
        q.ev.resize(evc[i]);
        q.lval.resize(2 * S);
        q.info.resize(sm.c);
        HIPCHK(ctx, hipMemcpy(q.lcnt.data(), (uint8_t *)ctx->b_lcnt.p + im.goff * 2, 2 * sm.G, hipMemcpyDeviceToHost));
        HIPCHK(ctx, hipMemcpy(q.lfull.data(), (uint16_t *)ctx->b_lfull.p + im.goff * 2 * EDDY_MAX_MP,
                              sizeof(uint16_t) * 2 * sm.G * EDDY_MAX_MP, hipMemcpyDeviceToHost));
        HIPCHK(ctx, hipMemcpy(q.rep.data(), (int *)ctx->b_rep.p + im.soff, sizeof(int) * S, hipMemcpyDeviceToHost));
        if (evc[i])
          HIPCHK(ctx, hipMemcpy(q.ev.data(), (Event *)ctx->b_ev.p + im.evoff, sizeof(Event) * evc[i],
                                hipMemcpyDeviceToHost));
        HIPCHK(ctx, hipMemcpy(q.lval.data(), (double *)ctx->b_lval.p + im.soff * 2, sizeof(double) * 2 * S,
                              hipMemcpyDeviceToHost));
        if (sm.c)
          HIPCHK(ctx, hipMemcpy(q.info.data(), (uint32_t *)ctx->b_info.p + im.infooff, sizeof(uint32_t) * sm.c,
                                hipMemcpyDeviceToHost));
        q.famkey.resize(fc[i]);
        q.famscore.resize(2 * (size_t)fc[i]);
        if (fc[i]) {
          HIPCHK(ctx, hipMemcpy(q.famkey.data(), (uint64_t *)ctx->b_famkey.p + im.foff, sizeof(uint64_t) * fc[i],
                                hipMemcpyDeviceToHost));
          HIPCHK(ctx, hipMemcpy(q.famscore.data(), (double *)ctx->b_famscore.p + im.foff * 2,
                                sizeof(double) * 2 * fc[i], hipMemcpyDeviceToHost));
        }
      }
    }
    return EDDY_OK;
  }
  ctx->err = "event/family capacity or hash collisions persisted after retries";
  return EDDY_ERR_CAPACITY;
}

void add_item(Chunk &ch, const std::vector<SetMeta> &sets, int set, int perm, const std::vector<uint32_t> &labbits,
              int N) {
  const SetMeta &sm = sets[set];
  ItemMeta im{};
  im.set = set;
  im.perm = perm;
  im.infooff = ch.pairs;
  im.laboff = (long long)ch.labels.size();
  im.soff = ch.sslots;
  im.goff = ch.gslots;
  ch.labels.insert(ch.labels.end(), labbits.begin(), labbits.end());
  int item = (int)ch.items.size();
  int nb = (sm.G + 31) / 32;
  for (int bi = 0; bi < nb; bi++)
    for (int bj = 0; bj <= bi; bj++) ch.tiles.push_back(make_uint2((unsigned)item, (unsigned)(bi << 16 | bj)));
  ch.items.push_back(im);
  ch.pairs += sm.c;
  ch.sslots += N + 2;
  ch.gslots += sm.G;
}

std::vector<uint32_t> label_bits(const std::vector<uint8_t> &cls) {
  std::vector<uint32_t> b((cls.size() + 31) / 32, 0);
  for (size_t s = 0; s < cls.size(); s++)
    if (cls[s]) b[s >> 5] |= 1u << (s & 31);
  return b;
}

// class labels of one permutation from its accepted order (main.cu:561-570)
std::vector<uint8_t> perm_labels(const std::vector<int> &order, int n1) {
  std::vector<uint8_t> cls(order.size());
  for (size_t i = 0; i < order.size(); i++) cls[order[i]] = (int)i < n1 ? 0 : 1;
  return cls;
}

int prepare_sets(eddy_ctx *ctx, const eddy_set_desc *sets, int nsets, const eddy_params *p,
                 std::vector<SetMeta> &meta) {
  const int N = ctx->N, NP = (N + 63) & ~63;
  meta.resize(nsets);
  std::vector<int> rows;
  std::vector<long long> rowoff;
  std::vector<uint32_t> prior;
  long long xoff = 0;
  for (int s = 0; s < nsets; s++) {
    const eddy_set_desc &d = sets[s];
    if (d.n_genes < 0 || d.n_genes > EDDY_MAX_G || (d.n_genes && !d.genes) || d.n_prior < 0 ||
        (d.n_prior && !d.prior)) {
      ctx->err = "invalid gene set (at most 1024 genes)";
      return EDDY_ERR_ARG;
    }
    SetMeta &m = meta[s];
    m.G = d.n_genes;
    m.c = d.n_genes * (d.n_genes - 1) / 2;
    m.xoff = xoff;
    m.poff = (long long)prior.size();
    edge_bounds(p, m.c, m.eb);
    for (int i = 0; i < d.n_genes; i++) {
      if (d.genes[i] < 0 || d.genes[i] >= ctx->genes) {
        ctx->err = "gene index out of range";
        return EDDY_ERR_ARG;
      }
      rows.push_back(d.genes[i]);
      rowoff.push_back(xoff + (long long)i * NP);
    }
    xoff += (long long)d.n_genes * NP;
    prior.resize(prior.size() + (m.c + 31) / 32 + 1, 0);
    for (int q = 0; q < d.n_prior; q++) {
      int x = d.prior[2 * q], y = d.prior[2 * q + 1];
      if (x < 0 || y < 0 || x >= d.n_genes || y >= d.n_genes) {
        ctx->err = "prior pair out of range";
        return EDDY_ERR_ARG;
      }
      if (x == y) continue;
      int t = std::max(x, y), c = std::min(x, y);
      long long pidx = (long long)t * (t - 1) / 2 + c;
      prior[m.poff + (pidx >> 5)] |= 1u << (pidx & 31);
    }
  }
  GROW(b_sets, sizeof(SetMeta) * nsets);
  GROW(b_xs, xoff);
  GROW(b_prior, sizeof(uint32_t) * prior.size());
  GROW(b_genes, sizeof(int) * rows.size());
  GROW(b_rowoff, sizeof(long long) * rowoff.size());
  hipStream_t st = ctx->stream;
  HIPCHK(ctx, hipMemcpyAsync(ctx->b_sets.p, meta.data(), sizeof(SetMeta) * nsets, hipMemcpyHostToDevice, st));
  if (!prior.empty())
    HIPCHK(ctx, hipMemcpyAsync(ctx->b_prior.p, prior.data(), sizeof(uint32_t) * prior.size(), hipMemcpyHostToDevice, st));
  if (!rows.empty()) {
    HIPCHK(ctx, hipMemcpyAsync(ctx->b_genes.p, rows.data(), sizeof(int) * rows.size(), hipMemcpyHostToDevice, st));
    HIPCHK(ctx, hipMemcpyAsync(ctx->b_rowoff.p, rowoff.data(), sizeof(long long) * rowoff.size(), hipMemcpyHostToDevice,
                               st));
    int nrows = (int)rows.size();
    k_gather<<<nrows, 64, 0, st>>>(ctx->dX, N, NP, (const int *)ctx->b_genes.p, (const long long *)ctx->b_rowoff.p,
                                   nrows, (uint8_t *)ctx->b_xs.p);
    HIPCHK(ctx, hipGetLastError());
  }
  HIPCHK(ctx, hipStreamSynchronize(st));
  return EDDY_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" int eddy_run_sets(eddy_ctx *ctx, const eddy_set_desc *sets, int nsets, const eddy_params *p,
                             eddy_set_result *results) {
  if (!ctx || (nsets && (!sets || !results)) || nsets < 0) return EDDY_ERR_ARG;
  if (!ctx->dX) return EDDY_ERR_STATE;
  CHK(check_params(ctx, p));
  hipSetDevice(ctx->device);
  ctx->timing = eddy_timing{};
  auto t0 = std::chrono::steady_clock::now();
  const int N = ctx->N, perms = p->perms, S = N + 2;
  std::vector<SetMeta> meta;
  CHK(prepare_sets(ctx, sets, nsets, p, meta));
  GROW(b_js, sizeof(double) * (size_t)std::max(1, nsets) * perms);
  HIPCHK(ctx, hipMemsetAsync(ctx->b_js.p, 0, sizeof(double) * (size_t)std::max(1, nsets) * perms, ctx->stream));
  GROW(b_counters, sizeof(unsigned long long) * 8);
  HIPCHK(ctx, hipMemsetAsync(ctx->b_counters.p, 0, sizeof(unsigned long long) * 8, ctx->stream));
  ctx->last_perms = perms;
  ctx->last_G.assign(nsets, 0);
  ctx->last_js.assign(nsets, std::vector<double>());
  ctx->last_p0.assign(nsets, Perm0());
  std::vector<long long> tests_eval(nsets, 0);
  std::vector<int> nan0(nsets, 0);
  std::vector<int> ran(nsets, 1);
  const std::vector<uint32_t> truebits = label_bits(ctx->classOf);
  const bool batched = p->mode == EDDY_MODE_BATCHED;
  // budget of pairs per chunk (device memory ~ 40 B/pair + per-item scratch)
  const long long pair_budget = 64LL << 20;

  auto flush = [&](Chunk &ch, bool capture) -> int {
    std::vector<long long> t(ch.items.size(), 0);
    CHK(run_chunk(ctx, meta, ch, p, perms, capture, &ctx->last_p0, t));
    for (size_t i = 0; i < ch.items.size(); i++) tests_eval[ch.items[i].set] += t[i];
    ch = Chunk();
    return EDDY_OK;
  };

  std::vector<int> order;
  if (batched) {
    // phase 1: permutation 0 (true labels) of every set
    Chunk ch;
    for (int s = 0; s < nsets; s++) {
      ctx->last_G[s] = meta[s].G;
      if (meta[s].G < 2) {  // a single (empty) network: likelihoods are NaN (main.cu:1345-1397)
        nan0[s] = 1;
        continue;
      }
      add_item(ch, meta, s, 0, truebits, N);
      if (ch.pairs > pair_budget) CHK(flush(ch, true));
    }
    CHK(flush(ch, true));
    std::vector<double> js0(nsets);
    for (int s = 0; s < nsets; s++) {
      if (nan0[s]) continue;
      double v;
      HIPCHK(ctx, hipMemcpy(&v, (double *)ctx->b_js.p + (size_t)s * perms, sizeof(double), hipMemcpyDeviceToHost));
      if (std::isnan(v)) nan0[s] = 1;
    }
    // phase 2: a producer thread replays the stream in the reference's order
    // (sets in order; a NaN-at-permutation-0 set consumes one permutation's
    // draws) while this thread batches the items and runs the GPU chunks
    std::vector<std::vector<uint32_t>> setlab(nsets);
    std::atomic<int> ready{0};
    double stream_ms = 0.0;
    std::thread producer([&]() {
      auto h0 = std::chrono::steady_clock::now();
      std::vector<int> ord;
      std::vector<uint8_t> cls(N);
      const int words = (N + 31) / 32;
      for (int s = 0; s < nsets; s++) {
        ctx->stream_gen.permutation(N, ord);  // permutation 0's draws
        if (!nan0[s]) {
          std::vector<uint32_t> &L = setlab[s];
          L.assign((size_t)(perms - 1) * words, 0u);
          for (int pm = 1; pm < perms; pm++) {
            ctx->stream_gen.permutation(N, ord);
            uint32_t *w = &L[(size_t)(pm - 1) * words];
            for (int i = ctx->n1; i < N; i++) w[ord[i] >> 5] |= 1u << (ord[i] & 31);  // class 2 (main.cu:561-570)
          }
        }
        ready.store(s + 1, std::memory_order_release);
      }
      stream_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    });
    int rc2 = EDDY_OK;
    const int words = (N + 31) / 32;
    std::vector<uint32_t> one(words);
    for (int s = 0; s < nsets && rc2 == EDDY_OK; s++) {
      while (ready.load(std::memory_order_acquire) <= s) std::this_thread::yield();
      if (nan0[s]) continue;
      ran[s] = perms;
      for (int pm = 1; pm < perms; pm++) {
        std::copy(setlab[s].begin() + (size_t)(pm - 1) * words, setlab[s].begin() + (size_t)pm * words, one.begin());
        add_item(ch, meta, s, pm, one, N);
      }
      std::vector<uint32_t>().swap(setlab[s]);
      if (ch.pairs > pair_budget) rc2 = flush(ch, false);
    }
    producer.join();
    ctx->timing.host_stream_ms += stream_ms;
    CHK(rc2);
    CHK(flush(ch, false));
  } else {
    // plain: one set and one permutation at a time, like main.cu:538-1460
    for (int s = 0; s < nsets; s++) {
      ctx->last_G[s] = meta[s].G;
      for (int pm = 0; pm < perms; pm++) {
        ctx->stream_gen.permutation(N, order);
        if (meta[s].G < 2) {
          nan0[s] = 1;
          break;
        }
        Chunk ch;
        add_item(ch, meta, s, pm, pm == 0 ? truebits : label_bits(perm_labels(order, ctx->n1)), N);
        CHK(flush(ch, pm == 0));
        if (pm == 0) {
          double v;
          HIPCHK(ctx, hipMemcpy(&v, (double *)ctx->b_js.p + (size_t)s * perms, sizeof(double), hipMemcpyDeviceToHost));
          if (std::isnan(v)) {
            nan0[s] = 1;
            break;
          }
        }
        ran[s] = pm + 1;
      }
    }
  }
  // p-values on the GPU
  GROW(b_nan0, sizeof(int) * std::max(1, nsets));
  GROW(b_pout, sizeof(double) * 6 * std::max(1, nsets));
  std::vector<double> pout((size_t)6 * nsets);
  if (nsets) {
    HIPCHK(ctx, hipMemcpyAsync(ctx->b_nan0.p, nan0.data(), sizeof(int) * nsets, hipMemcpyHostToDevice, ctx->stream));
    hipEvent_t q0, q1;
    hipEventCreate(&q0);
    hipEventCreate(&q1);
    hipEventRecord(q0, ctx->stream);
    k_pvalue<<<(nsets + 63) / 64, 64, 0, ctx->stream>>>((const double *)ctx->b_js.p, nsets, perms,
                                                         (const int *)ctx->b_nan0.p, (double *)ctx->b_pout.p);
    hipEventRecord(q1, ctx->stream);
    hipEventSynchronize(q1);
    float qms = 0;
    hipEventElapsedTime(&qms, q0, q1);
    ctx->timing.kernel_ms[4] += qms;
    ctx->timing.launches[4] += 1;
    hipEventDestroy(q0);
    hipEventDestroy(q1);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipMemcpyAsync(pout.data(), ctx->b_pout.p, sizeof(double) * 6 * nsets, hipMemcpyDeviceToHost,
                               ctx->stream));
  }
  std::vector<double> alljs((size_t)nsets * perms);
  if (nsets)
    HIPCHK(ctx, hipMemcpyAsync(alljs.data(), ctx->b_js.p, sizeof(double) * alljs.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (int s = 0; s < nsets; s++) {
    eddy_set_result &r = results[s];
    const double *o = &pout[(size_t)6 * s];
    r.status = (int)o[0];
    r.perms_run = nan0[s] ? 1 : ran[s];
    r.U0 = ctx->last_p0[s].U;
    r.js0 = nan0[s] ? -999.0 : alljs[(size_t)s * perms];
    r.pDDN = o[2];
    r.alpha = o[3];
    r.beta = o[4];
    long long c = meta[s].c;
    r.tests_ref = (long long)r.perms_run * S * c;
    r.tests_eval = tests_eval[s];
    ctx->last_js[s].assign(alljs.begin() + (size_t)s * perms, alljs.begin() + (size_t)(s + 1) * perms);
    if (nan0[s]) ctx->last_js[s][0] = -999.0;
  }
  unsigned long long cnt[8];
  HIPCHK(ctx, hipMemcpy(cnt, ctx->b_counters.p, sizeof(cnt), hipMemcpyDeviceToHost));
  for (int k = 0; k < 4; k++) ctx->timing.sched[k] = (long long)cnt[k];
  ctx->timing.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return EDDY_OK;
}

extern "C" int eddy_get_js(eddy_ctx *ctx, int set, double *out, int cap) {
  if (!ctx || set < 0 || set >= (int)ctx->last_js.size() || !out) return EDDY_ERR_ARG;
  const auto &v = ctx->last_js[set];
  if (cap < (int)v.size()) return EDDY_ERR_CAPACITY;
  std::copy(v.begin(), v.end(), out);
  return EDDY_OK;
}

// Assemble the capped CSR of permutation 0's unique networks from the
// compact capture (full rows + differing rows), in network order
// (structureUnique, routines.cu:63-115).
extern "C" int eddy_get_perm0_networks(eddy_ctx *ctx, int set, int *ids, int *nodes, int nodes_cap, int *epn,
                                       int *edges, int edges_cap, double *lval, int *n_edges) {
  if (!ctx || set < 0 || set >= (int)ctx->last_p0.size()) return EDDY_ERR_ARG;
  const Perm0 &q = ctx->last_p0[set];
  const int G = ctx->last_G[set], S = ctx->N + 2, n1 = ctx->n1;
  if (q.rep.empty()) return EDDY_ERR_STATE;
  // events by network, sorted by row
  std::vector<std::vector<const Event *>> evs(S);
  for (const Event &e : q.ev) evs[e.s].push_back(&e);
  for (auto &v : evs) std::sort(v.begin(), v.end(), [](const Event *x, const Event *y) { return x->t < y->t; });
  int u = 0, ne = 0;
  if (epn) epn[0] = 0;
  for (int s = 0; s < S; s++) {
    if (q.rep[s] != s) continue;
    int g = s <= n1 ? 0 : 1;
    size_t k = 0;
    int acc = 0;
    for (int t = 0; t < G; t++) {
      if (nodes) {
        if ((long long)u * G + t >= nodes_cap) return EDDY_ERR_CAPACITY;
        nodes[(size_t)u * G + t] = acc;
      }
      const uint16_t *par;
      int cnt;
      while (k < evs[s].size() && evs[s][k]->t < t) k++;
      if (k < evs[s].size() && evs[s][k]->t == t) {
        par = evs[s][k]->par;
        cnt = evs[s][k]->n;
      } else {
        par = &q.lfull[((size_t)g * G + t) * EDDY_MAX_MP];
        cnt = q.lcnt[(size_t)g * G + t];
      }
      for (int i = 0; i < cnt; i++) {
        if (edges) {
          if (ne >= edges_cap) return EDDY_ERR_CAPACITY;
          edges[ne] = par[i];
        }
        ne++;
      }
      acc += cnt;
    }
    if (ids) ids[u] = s;
    if (lval) {
      lval[u] = q.lval[u];
      lval[q.U + u] = q.lval[S + u];
    }
    u++;
    if (epn) epn[u] = ne;
  }
  if (n_edges) *n_edges = ne;
  return EDDY_OK;
}

extern "C" int eddy_get_edgelist(eddy_ctx *ctx, int set, int *nodes, int *epn, int *edges, int edges_cap) {
  if (!ctx || set < 0 || set >= (int)ctx->last_p0.size()) return EDDY_ERR_ARG;
  const Perm0 &q = ctx->last_p0[set];
  const int G = ctx->last_G[set];
  if (q.rep.empty() && G >= 2) return EDDY_ERR_STATE;
  int ne = 0;
  if (epn) epn[0] = 0;
  for (int g = 0; g < 2; g++) {
    int acc = 0;
    long long pidx = 0;
    for (int t = 0; t < G; t++) {
      if (nodes) nodes[g * G + t] = acc;
      for (int col = 0; col < t; col++, pidx++)
        if ((q.info[pidx] >> g) & 1) {
          if (edges) {
            if (ne >= edges_cap) return EDDY_ERR_CAPACITY;
            edges[ne] = col;
          }
          ne++;
          acc++;
        }
    }
    if (epn) epn[g + 1] = ne;
  }
  return EDDY_OK;
}

extern "C" int eddy_get_timing(eddy_ctx *ctx, eddy_timing *t) {
  if (!ctx || !t) return EDDY_ERR_ARG;
  *t = ctx->timing;
  return EDDY_OK;
}

// One permutation with caller-given labels; outputs in the reference's
// layouts (out23, the compacted unique CSR, out5 is not kept: BDeu per family
// is folded into the network sums, so out5 is reconstructed per unique
// network from the same family scores).
extern "C" int eddy_run_perm_detail(eddy_ctx *ctx, const eddy_set_desc *set, const uint8_t *labels,
                                    const eddy_params *p, eddy_perm_detail *out) {
  if (!ctx || !set || !labels || !out) return EDDY_ERR_ARG;
  if (!ctx->dX) return EDDY_ERR_STATE;
  CHK(check_params(ctx, p));
  hipSetDevice(ctx->device);
  const int N = ctx->N, S = N + 2;
  int n1 = 0;
  for (int s = 0; s < N; s++) n1 += labels[s] == 0;
  if (n1 != ctx->n1) {
    ctx->err = "labels must keep the class sizes";
    return EDDY_ERR_ARG;
  }
  std::vector<SetMeta> meta;
  CHK(prepare_sets(ctx, set, 1, p, meta));
  GROW(b_js, sizeof(double) * p->perms);
  HIPCHK(ctx, hipMemsetAsync(ctx->b_js.p, 0, sizeof(double) * p->perms, ctx->stream));
  GROW(b_counters, sizeof(unsigned long long) * 8);
  HIPCHK(ctx, hipMemsetAsync(ctx->b_counters.p, 0, sizeof(unsigned long long) * 8, ctx->stream));
  ctx->last_p0.assign(1, Perm0());
  ctx->last_G.assign(1, meta[0].G);
  ctx->last_perms = p->perms;
  const int G = meta[0].G;
  if (G < 2) {
    ctx->err = "detail run needs at least two genes";
    return EDDY_ERR_ARG;
  }
  Chunk ch;
  std::vector<uint8_t> cls(labels, labels + N);
  add_item(ch, meta, 0, 0, label_bits(cls), N);
  std::vector<long long> t(1);
  CHK(run_chunk(ctx, meta, ch, p, p->perms, true, &ctx->last_p0, t));
  double js;
  HIPCHK(ctx, hipMemcpy(&js, ctx->b_js.p, sizeof(double), hipMemcpyDeviceToHost));
  const Perm0 &q = ctx->last_p0[0];
  out->js = js;
  out->U = q.U;
  int ne = 0;
  CHK(eddy_get_perm0_networks(ctx, 0, out->uniq_ids, out->uni_nodes, S * G, out->uni_epn, out->uni_edges,
                              S * G * p->mp, out->lval, &ne));
  out->edSum = ne;
  if (out->bits) {
    // out23 from the per-pair decisions: network s of class g takes the
    // leave-one-out decision of the cell of its left-out sample
    // (kernel1.cu:85-89); this needs every leave-one-out decision
    if (p->schedule != EDDY_TESTS_ALL) {
      ctx->err = "bits output needs schedule EDDY_TESTS_ALL (every leave-one-out test)";
      return EDDY_ERR_ARG;
    }
    const int c = meta[0].c;
    std::vector<int> c1, c2;
    for (int s = 0; s < N; s++) (labels[s] ? c2 : c1).push_back(s);
    std::vector<uint8_t> xrow((size_t)G * N);
    for (int i = 0; i < G; i++) {
      std::vector<uint8_t> tmp((N + 63) & ~63);
      HIPCHK(ctx, hipMemcpy(tmp.data(), (uint8_t *)ctx->b_xs.p + (size_t)i * ((N + 63) & ~63), tmp.size(),
                            hipMemcpyDeviceToHost));
      std::copy(tmp.begin(), tmp.begin() + N, xrow.begin() + (size_t)i * N);
    }
    long long pidx = 0;
    for (int r = 1; r < G; r++)
      for (int col = 0; col < r; col++, pidx++) {
        uint32_t w = q.info[pidx];
        for (int s = 0; s < S; s++) {
          int g = s <= n1 ? 0 : 1;
          int k = g == 0 ? s - 1 : s - n1 - 1;
          int nk = g == 0 ? n1 : N - n1;
          int bit = (w >> g) & 1;
          if (k >= 0 && k < nk) {
            int sid = g == 0 ? c1[k] : c2[k];
            int a = xrow[(size_t)r * N + sid], b = xrow[(size_t)col * N + sid];
            if (a < 3 && b < 3) bit = (w >> (INFO_LOO_SHIFT(g) + a * 3 + b)) & 1;
          }
          out->bits[(size_t)s * c + pidx] = (uint8_t)bit;
        }
      }
  }
  if (out->out5) {
    // out5[(g*U + u)*G + t] = BDeu of row t's family of unique network u on
    // class g (run4's layout, kernel5.cu:237-257), from the family scores
    std::vector<int> ids(S), nodes((size_t)S * G), epn(S + 1), edges((size_t)S * G * p->mp);
    int n_e = 0;
    CHK(eddy_get_perm0_networks(ctx, 0, ids.data(), nodes.data(), S * G, epn.data(), edges.data(),
                                (int)edges.size(), nullptr, &n_e));
    const int U = q.U;
    for (int u = 0; u < U; u++)
      for (int t = 0; t < G; t++) {
        int b = epn[u] + nodes[(size_t)u * G + t];
        int e = (t == G - 1) ? epn[u + 1] : epn[u] + nodes[(size_t)u * G + t + 1];
        int np = std::min(e - b, 3);
        uint64_t key = (uint64_t)t | ((uint64_t)np << 10);
        for (int i = 0; i < np; i++) key |= (uint64_t)edges[b + i] << (12 + 10 * i);
        key += 1;
        int f = -1;
        for (size_t k = 0; k < q.famkey.size(); k++)
          if (q.famkey[k] == key) {
            f = (int)k;
            break;
          }
        if (f < 0) {
          ctx->err = "internal: family of a unique network row not scored";
          return EDDY_ERR_STATE;
        }
        for (int g = 0; g < 2; g++) out->out5[((size_t)g * U + u) * G + t] = q.famscore[(size_t)f * 2 + g];
      }
  }
  return EDDY_OK;
}

// Host build of the engine's edge decision (edgetest.h), for the tests that
// compare it with the oracle on many tables.  c: Bonferroni factor.
extern "C" int eddy_edge_decision_host(const int *tally9, int c, int prior, const eddy_params *p) {
  if (!tally9 || !p) return -1;
  eddy_edge_params ep;
  eddy_ctx tmp;
  tmp.lg15 = std::lgamma(1.5);
  tmp.lg2 = std::lgamma(2.0);
  tmp.lg3 = std::lgamma(3.0);
  edge_params(&tmp, p, ep);
  eddy_edge_bounds eb;
  edge_bounds(p, c, eb);
  return eddy_edge_decision(tally9, prior, &ep, &eb, (double)c);
}
