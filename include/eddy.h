/*
 * eddy.h - C ABI of libeddy: the MI355X (gfx950) EDDY engine.
 *
 * The reference (skimlab/eddy-gpu) has no library boundary: main.cu drives
 * six CUDA kernels through raw device pointers (prototypes at
 * definitions.cuh:78-136) and blocking copies, one permutation at a time.
 * This header replaces that internal operator API with a context-based one:
 *
 *   reference                                   libeddy
 *   ------------------------------------------  ---------------------------
 *   main.cu:27-56  device probe, default stream  eddy_ctx_create / _destroy
 *   main.cu:684-719 per-permutation cudaMalloc    dataset uploaded once
 *   main.cu:538-573 rand() permutation stream     eddy_stream_* (own TYPE_3
 *                                                 generator, same sequence)
 *   main.cu:575-1460 permutation loop, kernels    eddy_run_sets: all sets,
 *     run2/run2Scalable (kernel1.cu:389-477),     permutations and networks
 *     edgePerNetworkKernel (kernel2.cu:114),      of a batch on the GPU
 *     run22 (kernel3.cu:134), run25 (kernel4.cu:72),
 *     run4 (kernel5.cu:237), likelihood + JS
 *   main.cu:1462-1502 beta fit + Boost ibetac     computed on the GPU,
 *                                                 returned in eddy_set_result
 *   main.cu:782-819, 1085-1195, 1530-1641         eddy_get_perm0_networks,
 *     permutation-0 captures + EdgeList pass      eddy_get_edgelist
 *
 * Conventions: every function returns an eddy_status; nothing exits the
 * process.  The caller owns all host buffers; the context owns device memory
 * and one HIP stream.  A context is used by one host thread at a time.
 */
#ifndef EDDY_H
#define EDDY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  EDDY_OK = 0,
  EDDY_ERR_ARG = 1,      /* invalid argument */
  EDDY_ERR_HIP = 2,      /* HIP runtime error (message in eddy_last_error) */
  EDDY_ERR_CAPACITY = 3, /* a caller buffer is too small */
  EDDY_ERR_STATE = 4,    /* call out of order (no dataset, no results) */
  EDDY_ERR_NOMEM = 5,
  EDDY_ERR_NODEV = 6     /* no HIP device / libeddy built without one */
} eddy_status;

/* Per-set outcome, mirrors the three branches of main.cu:1465-1502 */
typedef enum {
  EDDY_SET_OK = 0,          /* results.txt row "%s\t%f\t%f\t%d" */
  EDDY_SET_FAILED_NAN = 1,  /* "failed with %d genes- one of the js scores was nan" */
  EDDY_SET_DOMAIN = 2       /* beta fit outside ibetac's domain (the reference's
                               Boost call would throw); pDDN is NaN */
} eddy_set_status;

typedef enum {
  EDDY_MODE_PLAIN = 0,   /* one gene set and one permutation per launch
                            sequence, as the reference's loop is structured */
  EDDY_MODE_BATCHED = 1  /* all sets x permutations of a batch in chunks of
                            thousands of items, each chunk's kernels one
                            HIP graph launch (eddy_launch_cfg.graph) */
} eddy_mode;

typedef enum {
  EDDY_TESTS_ALL = 0,    /* every leave-one-out test of every pair, as the
                            reference evaluates every (network, pair) */
  EDDY_TESTS_CUTOFF = 1  /* leave-one-out tests of a row only up to its mp-th
                            edge that no leave-one-out test can remove; later
                            columns cannot change any capped network */
} eddy_schedule;

typedef struct {
  int mp;                /* parent cap, -mp (main.cu:158-163): any positive
                            value, as in the reference (main.cu:170-172) */
  int perms;             /* r, -r (main.cu:165-168) */
  int p_adjust;          /* Bonferroni (1) or -rawp (0) (kernel1.cu:421-424) */
  double alpha_edge;     /* -pE, or from -t/-l (main.cu:180-186) */
  double alpha_edge_prior; /* 1-(1-aE)*pw^(1/l) (main.cu:188-189) */
  double alpha_ddn;      /* -pD (main.cu:150-153) */
  int mode;              /* eddy_mode */
  int schedule;          /* eddy_schedule */
  int learned;           /* 1: learned proposal windows the cutoff schedule's full tests (needs a
                            trained network); 3: no full tests before the plan asks for them */
} eddy_params;

/* One gene set after name lookup (main.cu:400-430) and prior loading
 * (main.cu:433-503). */
typedef struct {
  int n_genes;          /* G */
  const int *genes;     /* [G] dataset row of each set gene, set order */
  int n_prior;          /* number of prior pairs */
  const int *prior;     /* [2*n_prior] set positions (a, b); symmetric */
} eddy_set_desc;

typedef struct {
  int status;           /* eddy_set_status */
  int perms_run;        /* r, or 1 when permutation 0 is NaN */
  int U0;               /* unique networks of permutation 0 */
  double js0;           /* JS of the true labels (bits) */
  double pDDN;
  double alpha, beta;   /* method-of-moments beta fit */
  long long tests_ref;  /* reference-equivalent CI tests: sum over perms S*c */
  long long tests_eval; /* CI tests this engine evaluated */
} eddy_set_result;

typedef struct eddy_ctx eddy_ctx;

/* ---- context -------------------------------------------------------------- */
/* device -1 creates a host-only context (permutation stream only). */
int eddy_ctx_create(int device, eddy_ctx **out);
int eddy_ctx_destroy(eddy_ctx *ctx);
const char *eddy_last_error(const eddy_ctx *ctx);
const char *eddy_version(void);

/* Expression data: [genes x samples] row-major states 0,1,2 (reference values
 * -1,0,1; any other reference value is stored as 3 and never counted, F1 of
 * kernel1.cu:91-118).  class_of: [samples] true labels, 0 = class 1 (first
 * label of the class file), 1 = class 2 (main.cu:303-352).  At most 20000
 * samples, the reference's MAX_NUM_SAMPLES (definitions.cuh:17). */
int eddy_upload_dataset(eddy_ctx *ctx, const uint8_t *x, int genes, int samples,
                        const uint8_t *class_of);

/* ---- permutation stream (main.cu:538-573) ------------------------------- */
/* Reset to glibc's unseeded rand() state (the reference never calls srand). */
int eddy_stream_reset(eddy_ctx *ctx);
/* Standalone generator for tests: n successive rand() values. */
int eddy_stream_draw(eddy_ctx *ctx, int n, int32_t *out);
/* One permutation's draws (main.cu:541-570): for perm > 0 labels[samples]
 * receives the relabelling (0 = class 1); for perm 0 labels is untouched. */
int eddy_stream_permutation(eddy_ctx *ctx, int samples, int n1, int perm, uint8_t *labels);
/* The next `count` permutations' draws at once, as `count` calls of
 * eddy_stream_permutation would make them, through the stream index that
 * eddy_run_sets uses (worker threads scan the stream ahead).  c2 receives the
 * class-2 bits (1 = class 2) of each permutation: count * ceil(samples / 32)
 * words, sample i at bit i % 32 of word i / 32. */
int eddy_stream_class2_batch(eddy_ctx *ctx, int samples, int n1, int count, uint32_t *c2);

/* ---- the hot path ----------------------------------------------------------
 * Runs every set's permutation loop (main.cu:538-1460) and p-value
 * (main.cu:1462-1502) on the GPU, consuming the context's permutation stream
 * in the reference's order (sets in array order; a set whose permutation 0 is
 * NaN consumes one permutation's draws).  results: [nsets]. */
int eddy_run_sets(eddy_ctx *ctx, const eddy_set_desc *sets, int nsets,
                  const eddy_params *params, eddy_set_result *results);

/* JS values of set `set` of the last run: out[perms]. */
int eddy_get_js(eddy_ctx *ctx, int set, double *out, int cap);

/* Permutation-0 unique networks of a set of the last run (the `first_*`
 * captures of main.cu:1085-1195, 1406-1409): ids[U0] original network
 * indices, nodes[U0*G] capped CSR offsets, epn[U0+1], edges[epn[U0]],
 * lval[2*U0] normalised likelihoods (P1 || P2).  *n_edges receives epn[U0].
 * Any output may be NULL; caps are element counts. */
int eddy_get_perm0_networks(eddy_ctx *ctx, int set, int *ids, int *nodes,
                            int nodes_cap, int *epn, int *edges, int edges_cap,
                            double *lval, int *n_edges);

/* Uncapped parent lists of permutation 0's two full-data networks
 * (main.cu:1530-1612): nodes[2G], epn[3], edges[epn[2]]. */
int eddy_get_edgelist(eddy_ctx *ctx, int set, int *nodes, int *epn, int *edges,
                      int edges_cap);

/* ---- one permutation with explicit labels (parity tests) -----------------
 * labels[samples] in {0,1} (0 = class 1).  Outputs as in eddy_get_*; bits is
 * the reference's out23 [S*c] (network-major, S = samples+2); out5 is
 * [2*U*G] BDeu per (class, unique network, node).  Any output may be NULL. */
typedef struct {
  uint8_t *bits;    /* [S*c] */
  int *uniq_ids;    /* [S]   */
  int *uni_nodes;   /* [S*G] */
  int *uni_epn;     /* [S+1] */
  int *uni_edges;   /* [S*G*min(mp, G-1)] */
  double *out5;     /* [2*S*G] */
  double *lval;     /* [2*S] */
  int U, edSum;
  double js;
} eddy_perm_detail;

int eddy_run_perm_detail(eddy_ctx *ctx, const eddy_set_desc *set,
                         const uint8_t *labels, const eddy_params *params,
                         eddy_perm_detail *out);

/* eddy_run_perm_detail without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * device buffer of the call (the set's gathered rows and pooled tables, the
 * item's tallies, networks, events, families and captures) is carved from one
 * file-scope device array of EDDY_NOALLOC_SCRATCH_BYTES per GPU, which each
 * GPU the library runs on holds until the process ends; host staging is
 * pageable memory inside the context, and the kernels are launched one by
 * one (no graph objects).  The context's own buffers are untouched.  A set
 * and sample count whose buffers exceed the array is rejected with
 * EDDY_ERR_CAPACITY (a 300-gene set at 600 samples takes about 6 MiB,
 * 1024 genes at 20000 samples about 50 MiB).  Every call on one GPU shares
 * that GPU's array and finds the last call's contents, which it never relies
 * on; calls that overlap on one GPU (two contexts, two host threads) race.
 * Needs the dataset uploaded before (eddy_upload_dataset allocates). */
#define EDDY_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int eddy_run_perm_detail_noalloc(eddy_ctx *ctx, const eddy_set_desc *set,
                                 const uint8_t *labels, const eddy_params *params,
                                 eddy_perm_detail *out);

/* ---- timing of the last run (HIP events on the engine stream) ----------
 * kernel index: 0 count_test (tables on MFMA + tests), 1 loo (cutoff
 * schedule), 2 rows (networks + hashes), 3 score (dedup, BDeu, JS),
 * 4 pvalue, 5 gather. */
#define EDDY_NKERNELS 6
typedef struct {
  double total_ms;                   /* call entry to results on the host */
  double kernel_ms[EDDY_NKERNELS];   /* summed device time per kernel */
  long long launches[EDDY_NKERNELS];
  long long items;                   /* (set, permutation) items processed */
  long long pairs;                   /* sum over items of G(G-1)/2 */
  long long tests;                   /* chi-square tests evaluated */
  double host_stream_ms;             /* host time replaying the rand() stream */
  long long sched[4];                /* cutoff schedule: rows past their plan,
                                        columns tested past plans, planned
                                        (pair, class) entries, rows */
  long long stats[6];                /* BDeu histogram bytes read (child and
                                        parent states per sample, both
                                        classes), families scored, unique
                                        networks, events, histogram bin bytes
                                        (4 q r per family and class), 0 */
  double phase1_ms;                  /* host wall time before the first chunk
                                        (set preparation, features, windows) */
  double phase_ms[16];               /* workgroup time per phase summed over
                                        items: score 0 dedup, 1 confirm, 2
                                        unique list, 3 families, 4 BDeu, 5 sums,
                                        6 likelihood/JS; rows 7 setup, 8 rows;
                                        loo 10 plan, 11 packed tests, 12 cutoff
                                        check and fallback; count 13 staging
                                        and MFMA, 14 tests and stores */
  double split_ms[7];                /* summed device time per kernel of the
                                        batched sequence: 0 count, 1 loo,
                                        2 rows, 3 score, 4 BDeu, 5 tail,
                                        6 the class-1 gather before count */
  long long split_launches[7];
} eddy_timing;
int eddy_get_timing(eddy_ctx *ctx, eddy_timing *t);

/* ---- launch shapes and the autotuner (no counterpart in the reference,
 * whose launch shapes are fixed in main.cu:737-1290) ----------------------
 * Every field changes only how the work is laid on the GPU, never a result:
 * every configuration gives bit-identical outputs. */
typedef struct {
  int graph;        /* 1: each chunk's kernel sequence (clear, count, loo, rows,
                       score, BDeu, tail, timing events) is one hipGraph per
                       slot, instantiated once and updated in place for every
                       chunk; 0: the kernels are launched one by one */
  int tile_order;   /* count kernel: 0 = workgroup i takes tile i, 1 = each XCD
                       takes a contiguous run of the tile list */
  int item_order;   /* per-item kernels: 0 = largest set first, 1 = chunk order */
  int loo_waves;    /* loo kernel instance: 4 or 3 waves per SIMD */
  int bdeu_groups;  /* lane groups per wave in the BDeu of 2-3 parent families:
                       1, 2 or 4; 0 = from the mask length */
  int bdeu_blocks;  /* persistent workgroups of the BDeu kernel, 64..16384 */
} eddy_launch_cfg;
int eddy_get_launch_cfg(eddy_ctx *ctx, eddy_launch_cfg *cfg);
/* EDDY_ERR_ARG (and no change) for a value outside the ranges above */
int eddy_set_launch_cfg(eddy_ctx *ctx, const eddy_launch_cfg *cfg);

typedef struct {
  int from_cache;      /* 1: the configuration came from cache_path */
  int trials;          /* calibration launches timed */
  double tune_ms;      /* wall time of the call */
  double ms_default;   /* calibration chunk, kernels of the starting configuration */
  double ms_best;      /* the same chunk with the chosen configuration */
  long long items, pairs; /* the calibration chunk */
  double kernel_ms[6];  /* the chosen configuration on the calibration chunk (min of
                           3 serial runs): count (without the class-1 gather), loo,
                           rows, score, BDeu, tail */
} eddy_tune_report;
/* Picks the launch configuration for this dataset and these sets: one
 * calibration chunk (sets sampled across the list, permutations drawn from a
 * private generator, so the rand() stream is untouched) runs serially under
 * candidate configurations, knob by knob; each knob keeps the candidate whose
 * kernels' HIP-event time is at least 2% below the current choice.  The choice
 * is installed in the context (as eddy_set_launch_cfg) and written to *out.
 * cache_path (may be NULL): a text file of earlier choices keyed by device,
 * samples, classes, mp, schedule and largest set; a matching line is used
 * without calibrating, a new choice is appended.  Results of later runs do not
 * depend on the choice.  Does not touch the last run's retrievable results. */
int eddy_autotune(eddy_ctx *ctx, const eddy_set_desc *sets, int nsets, const eddy_params *params,
                  const char *cache_path, eddy_launch_cfg *out, eddy_tune_report *report);

/* Host build of the kernels' all-at-once leave-one-out check: 1 when every
 * leave-one-out tally of tally9 is certainly an edge, 0 when certainly none is,
 * -1 when the per-cell decisions are needed (tests compare it with the
 * oracle's decision of each leave-one-out tally). */
int eddy_loo_all_certain_host(const int *tally9, int c, int prior, const eddy_params *p);

/* Host build of the engine's own edge decision for one 3x3 tally
 * (tally9[r*3+k], r = row-gene state): 1 edge, 0 none.  Tests compare it
 * with the CPU oracle; it runs on the CPU and needs no context. */
int eddy_edge_decision_host(const int *tally9, int c, int prior, const eddy_params *p);
/* Leave-one-out decision mask of one 3x3 class tally (bit k: the tally minus
 * one count in cell k; empty cells copy `full`), as the kernels compute it
 * (kernel1.cu:236-419 tests each leave-one-out tally in full); host build,
 * for tests. */
int eddy_edge_loo_mask_host(const int *tally9, int c, int prior, int full, const eddy_params *p);
/* Chi-square and dof of a 3x3 tally: the reference's formula (fast = 0,
 * kernel1.cu:121-203), the kernels' reciprocal form (fast = 1), or the
 * kernels' leave-one-out form for the tally minus one count in cell k
 * (fast = 2 + k, k = 0..8; the cell must be nonempty); host build, for
 * tests. */
int eddy_chi2_host(const int *tally9, int fast, int *dof, double *chi2);
/* The count kernel's single-precision screen of the full test, host build,
 * for tests (no counterpart in the reference: it only decides a test early
 * where the reference's decision kernel1.cu:420-431 is already certain): 1
 * edge, 0 none, -1 when it leaves the decision to the double-precision path.
 * eddy_chi2_host(fast = 11) gives the screen's x (as chi-square, no Yates
 * correction). */
int eddy_screen32_host(const int *tally9, int c, int prior, const eddy_params *p);
/* The learned proposal's training order (no counterpart in the reference):
 * out[e * m + i] = the i-th of the m sampled rows (of n) that epoch e visits,
 * as the training kernels draw it for `seed`; host build, for tests. */
int eddy_mlp_order_host(uint64_t seed, long long n, long long m, int epochs, int *out);

/* ---- learned proposal ------------------------------------------------------
 * A 12 -> 64 -> 64 -> 1 network (ReLU, ReLU, sigmoid; FP32) scores each gene
 * pair from 12 features of its pooled table (Pearson chi2/n, log1p chi2,
 * mutual information, both marginal entropies, empty-cell fraction, smallest
 * expected count, n, dof, prior flag, Bonferroni factor, Cramer's V): the
 * probability that the pair is a full-test edge in a permuted class.  With
 * eddy_params.learned = 1 (batched, cutoff schedule) the engine sums the scores
 * along each row and runs the count kernel's full tests only up to the column
 * where they reach mp + margin; every later test the cutoff still needs is run
 * on demand, so results never depend on the network.  No counterpart in the
 * reference (whose tests run in a fixed order, kernel3.cu:70-109). */
#define EDDY_MLP_FEATURES 12
#define EDDY_MLP_NPARAMS 5057
int eddy_mlp_init(eddy_ctx *ctx, uint64_t seed);
int eddy_mlp_get_params(eddy_ctx *ctx, float *out, int n);
int eddy_mlp_set_params(eddy_ctx *ctx, const float *in, int n);
int eddy_mlp_set_margin(eddy_ctx *ctx, float kappa);
/* prob[n] = network output for X[n][12] */
int eddy_mlp_forward(eddy_ctx *ctx, const float *X, int n, float *prob);
/* mean binary cross-entropy gradient over (X[n][12], Y[n]) and its loss */
int eddy_mlp_gradient(eddy_ctx *ctx, const float *X, const float *Y, int n, float *grad, float *loss);
/* Adam on the GPU, `epochs` passes over shuffled mini-batches */
int eddy_mlp_train(eddy_ctx *ctx, const float *X, const float *Y, long long n, int epochs, int batch, float lr,
                   uint64_t seed, float *loss_out);
/* Runs `sets` without the network (every full test; leave-one-out tests up to
 * the exact cutoff) and returns one training row per gene pair: X[n][12] features, Y[n] fraction of (permutation > 0, class)
 * full-test edges.  *n_out = number of pairs (EDDY_ERR_CAPACITY if > cap). */
int eddy_learn_collect(eddy_ctx *ctx, const eddy_set_desc *sets, int nsets, const eddy_params *params, float *X,
                       float *Y, long long cap, long long *n_out);

/* The learned job: the first train_sets sets run without the network and
 * log the full-test outcomes of their first label_perms permutations (0: all);
 * the network is then trained on the GPU (Adam, `epochs` passes of
 * `batch`-row mini-batches over a uniform sample of max_rows logged rows (0:
 * all), learning rate lr, initialised from seed) on a second stream while the
 * next sets run with every full test deferred to the plan; sets launched after
 * the training finished use the network's windows (margin `margin`).
 * Results, networks and the permutation stream equal eddy_run_sets over all
 * sets, whichever sets use the windows. */
typedef struct {
  int train_sets;
  int epochs;
  int batch;
  float lr;
  float margin;
  uint64_t seed;
  long long max_rows;
  int label_perms;
} eddy_learn_cfg;

typedef struct {
  double collect_ms, train_ms, apply_ms; /* train_ms: GPU time of training (second stream) */
  long long train_rows;
  float loss_initial, loss_final;
  long long tests_collect, tests_apply;
  int trained;            /* 0 when too few rows were logged (< 64): no windows were used */
  int first_windowed_set; /* first set run with the network's windows, -1 if none */
} eddy_learn_report;

int eddy_run_learned(eddy_ctx *ctx, const eddy_set_desc *sets, int nsets, const eddy_params *params,
                     const eddy_learn_cfg *cfg, eddy_set_result *results, eddy_learn_report *report);

/* Re-permutation driver (no counterpart in the reference, which would be run
 * again by hand on the chosen sets): the `top` sets with the smallest pDDN in
 * `first` (status EDDY_SET_OK; ties by set index) run again with `perms`
 * permutations on a fresh permutation stream (as a new `eddy` process would
 * draw them), in set order.  chosen[k] receives the k-th re-run set's index,
 * results[k] its result, *nchosen their count (<= top).  The stream is left
 * after the re-run's draws. */
int eddy_run_repermute(eddy_ctx *ctx, const eddy_set_desc *sets, int nsets, const eddy_params *params,
                       const eddy_set_result *first, int top, int perms, int *chosen, eddy_set_result *results,
                       int *nchosen);

/* ---- synthetic workloads (seeded DAG data in the benchmark shapes) ------- */
typedef struct {
  uint64_t seed;
  int genes, nA, nB;
  double rewire;
  int nsets, gmin, gmax;
  int prior_fixed;        /* >= 0: fixed prior edges per set */
  double prior_per_gene;  /* else floor(prior_per_gene * G) */
  int mp, perms;
  double prior_weight;
} eddy_synth_cfg;

int eddy_synth_config(int k, eddy_synth_cfg *cfg);
int eddy_synth_dataset(uint64_t seed, int genes, int nA, int nB, double rewire,
                       uint8_t *x);
int eddy_synth_sets(uint64_t seed, int genes, int nsets, int gmin, int gmax,
                    int prior_fixed, double prior_per_gene, int *set_sizes,
                    int *set_genes, long long genes_cap, int *prior_counts,
                    int *prior_pairs, long long pairs_cap);

#ifdef __cplusplus
}
#endif

#endif /* EDDY_H */
