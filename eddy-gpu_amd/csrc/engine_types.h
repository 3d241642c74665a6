// engine_types.h - layouts shared by the host engine and the gfx950 kernels.
#pragma once
#include <stdint.h>

#include "edgetest.h"

// Engine limits.  The reference's: MAX_GENESET_SIZE 200 genes per set (a
// stack array, definitions.cuh:19, main.cu:239), MAX_NUM_SAMPLES 20000
// samples (definitions.cuh:17), any positive parent cap (main.cu:170-172).
#define EDDY_MAX_G 1024     // genes per set (10-bit family keys)
#define EDDY_MAX_N 20000    // samples (S = N + 2 networks; sample and network ids fit 16 bits)
#define EDDY_NARROW_MP 16   // parent caps up to this keep a row list in registers
#define EDDY_NWAVES 4       // waves per workgroup in the per-item kernels

// Per-pair info word written by the test kernels and read by the network
// kernel: class g in {0,1} (class 1 / class 2 of the reference) owns the 16
// bits at 16g, so each class's half is written by one thread with a plain
// 16-bit store.
//   bit 16g          full-data test of class g is an edge
//   bit 16g+1        the full test of class g was evaluated
//   bit 16g+2        the leave-one-out tests of class g were evaluated
//   bits 16g+3..+11  leave-one-out test per cell (r*3+k) of class g; a cell
//                    with no sample holds the full bit
#define INFO_FULL(g) (1u << (16 * (g)))
#define INFO_FEVAL(g) (1u << (16 * (g) + 1))
#define INFO_EVAL(g) (1u << (16 * (g) + 2))
#define INFO_LOO_SHIFT(g) (16 * (g) + 3)
#define INFO_FULL_BIT(w, g) (((w) >> (16 * (g))) & 1u)
// bits within one class's 16-bit half
#define IH_FULL 1u
#define IH_FEVAL 2u
#define IH_EVAL 4u
#define IH_LOO_SHIFT 3

struct SetMeta {
  int G;            // genes in the set
  int c;            // G(G-1)/2
  long long xoff;   // byte offset of the set's rows [G x N] in the gathered data
  long long poff;   // u32-word offset of the set's prior bits (one bit per pair)
  eddy_edge_bounds eb;  // x intervals where the edge decision is certain (c-dependent)
  long long pairbase;   // sum of c over the run's earlier sets (per-set pair arrays)
  long long genebase;   // sum of G over the run's earlier sets (per-set row arrays)
  int evhint;           // host only: expected events of a permuted item (0 = none yet)
};

struct ItemMeta {   // one (gene set, permutation)
  int set;
  int perm;
  long long infooff;  // pair offset into info / tables
  long long laboff;   // u32-word offset of the label bitmask (bit = class 2)
  long long evoff;    // event offset (capacity ev_cap per item)
  long long soff;     // network-indexed scratch offset (S per item)
  long long goff;     // gene-indexed scratch offset (G per item)
  long long foff;     // family scratch offset (fcap per item)
  int evcap;          // event capacity of this item
  int fcap;           // family capacity of this item
  long long fhoff;    // family hash table offset (fhsize entries)
  int fhsize;         // power of two >= 2 * fcap
  int nolog;          // a re-run: its first run already logged its label hits
  int capmin;         // host only: event capacity at least this (a re-run's exact count)
};

// one device-to-device copy of the permutation-0 capture
struct Slice {
  const void *src;
  void *dst;
  long long bytes;
};

// per-pair tallies in the cutoff schedule: class 1's 3x3 tally as packed u16
// counts; class 2's is the pair's pooled table (same layout, once per set)
// minus class 1's.  Cells 0..7 take two 8-byte words (8 u16).  Cell 8 is the
// class size minus the other eight when every sample of the dataset holds one
// of the three states (each sample then falls in exactly one cell); a dataset
// with samples outside them stores cell 8 in a third word (12 u16).
#define EDDY_TAB_DENSE 8
#define EDDY_TAB_FULL 12

// statistics counters (schedule, score statistics, per-phase workgroup
// time): EDDY_CNT_COPIES copies of EDDY_NCNT counters, one 256-byte line per
// copy, each workgroup adding into the copy of its index modulo the count
// (same-address atomics from every workgroup of a launch would serialize);
// the host sums the copies
#define EDDY_NCNT 32
#define EDDY_CNT_COPIES 64

struct KArgs {
  // dataset-level
  int N, n1, n2, S;
  int mp, p_adjust, plain;  // mp: the parent cap, at most the largest set's G - 1 (no row has more parents)
  int mpw;                 // u16 per row in lfull: mp
  int ht_size;             // network hash table entries (power of two >= 1.5 S)
  int score_shared;        // bytes of the score kernel's phase-shared LDS region
  int score_groups;        // lane groups per wave in the score kernel's BDeu of 2-3 parent families
  eddy_edge_params ep;     // thresholds, lgamma(p+1) constants, underflow-cut interval
  const double *lgtab;     // [82][lgstride]: lgamma(n + 4/d)
  int lgstride;
  uint64_t hseed;
  // sets
  const SetMeta *sets;
  const eddy_screen_thr *thr;  // per set: the full tests' screen thresholds for class 1's and class 2's sizes
  const uint8_t *xs;       // gathered set rows
  uint8_t *xc;             // [(goff + g) * kp + k]: each item's rows at its class-1 samples (cutoff schedule)
  int kp;                  // class-1 samples rounded up to 64
  const uint32_t *prior;   // prior bits
  // items
  const ItemMeta *items;
  const int *order;        // per-item kernels: workgroup b runs item order[b]
  const uint32_t *labels;
  // work lists
  const uint32_t *tiles;   // K1 work list: item<<16 | bi<<8 | bj
  // per-pair outputs
  uint32_t *info;
  uint16_t *tab;           // [pair][tabw] class 1's counts (cutoff mode)
  uint16_t *totals;        // [set pairbase + p][tabw]: pooled table of each pair over all samples
  int tabw;                // u16 per tally: EDDY_TAB_DENSE or EDDY_TAB_FULL
  // per-item scratch
  uint16_t *lfull;         // [goff*2 + g*G + t][mpw]
  uint16_t *lcnt;          // [goff*2 + g*G + t]
  uint8_t *sflags;         // [goff*2 + g*G + t] observed-state flags (bit v)
  int *evcount;            // [item]
  uint64_t *hnet;          // [soff + s]
  int *rep;                // [soff + s] first occurrence of each network
  uint16_t *clsg;          // [soff + k]: class-1 samples, then class-2 samples
  int *evorder;            // [evoff + e]
  uint32_t *evkey;         // [evoff + e]: s | t << 16 of event e (written with it by the rows kernel)
  uint16_t *evt;           // [evoff + k]: row of the k-th event in network order (score kernel)
  uint16_t *evpar;         // [(evoff + e) * evw]: parent count n, then the n parents of event e
  int evw;                 // u16 per event in evpar: mp + 1
  int *evfam;              // [evoff + e]
  uint64_t *famkey;        // [foff + f]
  unsigned long long *famht;  // [fhoff + i] family hash table keys
  int *famhv;              // [fhoff + i] family hash table values
  double *famscore;        // [(foff + f)*2 + g]
  double *lval;            // [soff*2 ...] 2 per network slot
  int *uniq_u;             // [item]
  int *famcount;           // [item] distinct families scored
  double *js;              // [set*perms + perm]
  int perms;
  int *status;             // [item] error bits
  const uint8_t *plan_extra;  // [goff*2 + g*G + t] extra full edges in the plan (learned), or null
  float *feat;             // [(pairbase + p) * 12] learned-proposal features (k_features)
  unsigned *labhits;       // [pairbase + p] full-test edges over permutations > 0 (training labels), or null
  int label_perms;         // > 0: permutations 1..label_perms run every full test (and log it) despite win
  const int *win;          // [genebase + t] last column whose full test the count kernel runs, or null
  unsigned long long *counters;  // statistics (optional): EDDY_CNT_COPIES x EDDY_NCNT
  uint64_t *bits;          // [((goff*2 + g*G + gene)*3 + v)*bitw + w]: class-g samples (class-list order) in state v
  int bitw;                // words per mask: ceil(max(n1, n2) / 64)
  int rows_maxg;           // largest set of the launch (row kernel LDS layout)
  long long *tests;        // [item] evaluated tests
  // score kernel -> BDeu kernel -> score tail kernel
  int2 *bd_tasks;          // (item, family * 2 + class) of every 2-3 parent family to score
  unsigned *bd_count;      // tasks listed (may exceed bd_cap: the rest were scored in place)
  int bd_cap;
  int *sc_ulist;           // [soff + u] unique networks in order
  int *sc_evst;            // [soff + s] first event of network s (network order)
  int *sc_famfull;         // [goff*2 + g*G + t] family of row t in class g's full network
  int tail_shared;         // bytes of the tail kernel's region for base scores, then KL terms
  uint32_t *fbits;         // [(goff*2 + g*G + t) * fbw + w]: full edges of row t, class g, columns 32w..32w+31
  int fbw, fbwp;           // words per row (ceil(largest G / 32)) and the power of two >= it
  int tile_order;          // count kernel: 1 = each XCD takes a contiguous run of the tile list
  // per-item global scratch of the score kernels when their network tables do
  // not fit LDS (sample counts in the thousands): item i owns bytes
  // [i * gscr_item, (i + 1) * gscr_item): the network hash table, then
  // rep / evst / work / ulist (score kernel), later the KL terms (tail)
  unsigned char *gscr;
  long long gscr_item;
};

enum {
  ST_EV_OVERFLOW = 1,
  ST_FAM_OVERFLOW = 2,
  ST_COLLISION = 4,
  ST_NOT_EVALUATED = 8,
};
