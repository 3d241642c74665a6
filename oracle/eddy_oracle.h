/*
 * eddy_oracle.h - CPU oracle for the EDDY permutation loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (skimlab/eddy-gpu main.cu, kernel1-5.cu, routines.cu), used by the
 * test suite and by the CPU-baseline leg of bench.py as the checker.  The
 * product (libeddy, the eddy CLI) never links, loads or calls it.
 *
 * It follows the reference's brute-force algorithm: a full leave-one-out 3x3
 * tally for every (network, pair), AS 147 for every test, pairwise network
 * comparison for dedup, brute-force BDeu counts, glibc rand() for the
 * permutation stream.  The only substitution is the FP64 log/exp used by the
 * chi-square p-value, the likelihoods and the JS divergence, which come from
 * fpmath.h (the same source the GPU kernels compile) so that both sides round
 * identically; BDeu uses glibc lgamma, like the lgamma table the GPU reads.
 *
 * Parity status per function is listed at the top of eddy_oracle.c.
 */
#ifndef EDDY_ORACLE_H
#define EDDY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int mp;                /* parent cap (main.cu:158-163) */
  int perms;             /* permutations r (main.cu:165-168) */
  int pAdjust;           /* Bonferroni on/off (-rawp, main.cu:141-142) */
  double alphaEdge;      /* main.cu:100-104, 180-186 */
  double alphaEdgePrior; /* 1-(1-aE)*pw^(1/lambda), main.cu:188-189 */
  double alphaDDN;       /* main.cu:150-153 */
  int threads;           /* OpenMP threads for the tally stage (<=1: serial) */
} eo_params;

/* Everything one permutation produces.  Buffers are caller-owned; any pointer
 * may be NULL to skip that output.  Bounds: S = n1+n2+2, c = G(G-1)/2.   */
typedef struct {
  uint8_t *edgebits; /* [S*c]   out23, network-major (kernel1.cu:389-477) */
  int *ppn;          /* [2G]    observed-state counts (kernel1.cu:338-387) */
  int *stf;          /* [6G]    observed-state flags */
  int *nodes;        /* [S*G]   capped CSR node offsets (kernel3.cu) */
  int *edges;        /* [S*(G-1)*min(mp,G)] capped parents */
  int *edgesPN;      /* [S+1]   edge prefix (main.cu:842-865) */
  int *uniqueN;      /* [S]     first-occurrence flags (main.cu:1060-1084) */
  int *uni_ids;      /* [S]     original ids of unique networks */
  int *uni_nodes;    /* [S*G]   compacted CSR (routines.cu:63-115) */
  int *uni_edges;    /* [edges bound] */
  int *uni_epn;      /* [S+1] */
  double *out5;      /* [2*S*G] BDeu per (class, unique net, node) */
  double *lval;      /* [2*S]   normalised likelihoods P1 || P2 */
  /* scalar outputs */
  int S, c, E, U, edSum;
  double js;         /* JS in bits (NaN if the likelihoods are NaN) */
} eo_perm_out;

/* H2: split the set's rows into class-ordered matrices (main.cu:575-605).
 * data: [numGenes x numSamples] int, classIDs: [numSamples] in {0,1}.
 * t1: [G x n1], t2: [G x n2]. */
void eo_class_split(const int *data, int numSamples, const int *setIdx, int G,
                    const int *classIDs, int n1, int n2, int *t1, int *t2);

/* H1: one permutation's label update from glibc rand() (main.cu:538-573).
 * Consumes rand() exactly like the reference; applies labels when perm > 0. */
void eo_perm_labels(int numSamples, int n1, int permNum, int *classIDs);

/* H3-H15 for one permutation given the class matrices. prior: [G*G] 0/1. */
void eo_run_perm(const int *t1, const int *t2, int G, int n1, int n2,
                 const int *prior, const eo_params *p, eo_perm_out *out);

/* H16: beta fit + ibetac (routines.cu:494-508, main.cu:1462-1502).
 * Returns pDDN; alpha and beta receive the fit; *domain_err is set when the
 * reference's Boost call would raise a domain error. */
double eo_pvalue(const double *js, int perms, double *alpha, double *beta,
                 int *domain_err);

/* Whole gene set (main.cu:359-1502) on already-parsed inputs.  Uses and
 * advances the process-wide rand() stream.  js_out: [perms].  Returns the
 * number of permutations run (1 when perm 0 is NaN: main.cu:1431-1435).
 * *failed = 1 when the reference would print the "failed" line. */
int eo_run_set(const int *data, int numSamples, const int *trueClass, int n1,
               int n2, const int *setIdx, int G, const int *prior,
               const eo_params *p, double *js_out, int *failed, double *pDDN,
               int *U0);

/* H17: uncapped parent lists (PARENTS_LIMIT = INT_MAX) of permutation 0's
 * two full-data networks for the EdgeList file (main.cu:1530-1612):
 * bits0 / bitsL are out23 rows of networks 0 and S-1 ([c] each); nodes[2G]
 * receives the CSR offsets, edges the parents (at most edges_cap), edgesPN
 * {0, E0, E0 + E1}.  Returns the edge count, or -1 if edges_cap is short. */
int eo_uncapped_pair(const uint8_t *bits0, const uint8_t *bitsL, int G, int *nodes, int *edges, int edges_cap,
                     int edgesPN[3]);

/* Reset the rand() stream to glibc's unseeded state (srand(1)). */
void eo_reset_stream(void);

/* Primitive hooks used by the oracle's own tests. */
double eo_gammds(double x, double p);
double eo_chi2(const int tally[9], int *dof);
int eo_edge_decision(const int tally[9], int c, int prior, const eo_params *p);
double eo_bdeu_family(const int *nodeData, const int *const *parentData,
                      int nParents, const int *nodeStates,
                      const int *const *parentStates, int samples);

/* FP64 primitives (fpmath.h) for the tests that pin them. */
double eo_exp(double x);
double eo_log(double x);
double eo_lgamma_pos(double x);
double eo_ibetac(double a, double b, double x, int *domain_err);

/* CLI restatement of main.cu:27-1711 (file outputs only). */
int eo_main(int argc, char **argv);

#ifdef __cplusplus
}
#endif

#endif
