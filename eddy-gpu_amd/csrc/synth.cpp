// synth.cpp - seeded synthetic EDDY workloads: a sparse linear-Gaussian DAG
// per class (class B rewires a fraction of class A's edges), tertile
// discretisation, random gene sets with random prior edges.
//
// Everything is bit-reproducible from the seed: std::mt19937_64 (its output
// sequence is fixed by the C++ standard), uniforms from the top 53 bits,
// normals from Box-Muller written here (std::*_distribution is not
// bit-portable across library versions).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "eddy.h"

namespace {

struct Rng {
  std::mt19937_64 g;
  explicit Rng(uint64_t s) : g(s) {}
  uint64_t next() { return g(); }
  double u01() { return (double)(g() >> 11) * 0x1p-53; }
  uint64_t below(uint64_t n) { return n ? g() % n : 0; }
  double normal() {
    double u1 = 1.0 - u01();  // (0, 1]
    double u2 = u01();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

struct Dag {
  std::vector<int> order;                  // topological order (gene ids)
  std::vector<std::vector<int>> parents;   // per gene
  std::vector<std::vector<double>> w;      // per gene, per parent
};

Dag make_dag(Rng &r, int G, int maxPar) {
  Dag d;
  d.order.resize(G);
  for (int i = 0; i < G; i++) d.order[i] = i;
  for (int i = G - 1; i > 0; i--) std::swap(d.order[i], d.order[r.below(i + 1)]);
  d.parents.assign(G, {});
  d.w.assign(G, {});
  for (int i = 0; i < G; i++) {
    int gene = d.order[i];
    int k = (int)r.below(maxPar + 1);
    if (k > i) k = i;
    while ((int)d.parents[gene].size() < k) {
      int p = d.order[r.below(i)];
      if (std::find(d.parents[gene].begin(), d.parents[gene].end(), p) == d.parents[gene].end()) {
        d.parents[gene].push_back(p);
        d.w[gene].push_back(0.5 + 0.5 * r.u01());
      }
    }
  }
  return d;
}

Dag rewire(Rng &r, const Dag &a, double frac) {
  Dag b = a;
  int G = (int)a.order.size();
  std::vector<int> posOf(G);
  for (int i = 0; i < G; i++) posOf[a.order[i]] = i;
  for (int i = 0; i < G; i++) {
    int gene = a.order[i];
    auto &ps = b.parents[gene];
    for (size_t e = 0; e < ps.size(); e++) {
      if (r.u01() >= frac) continue;
      if ((int)ps.size() >= posOf[gene]) continue;  // no free earlier gene
      for (int tries = 0; tries < 64; tries++) {
        int p = a.order[r.below(posOf[gene])];
        if (std::find(ps.begin(), ps.end(), p) == ps.end()) {
          ps[e] = p;
          break;
        }
      }
    }
  }
  return b;
}

void sample(Rng &r, const Dag &d, int n, int col0, int N, std::vector<double> &x) {
  int G = (int)d.order.size();
  for (int s = 0; s < n; s++) {
    for (int i = 0; i < G; i++) {
      int g = d.order[i];
      double v = 0.0;
      for (size_t e = 0; e < d.parents[g].size(); e++) v += d.w[g][e] * x[(size_t)d.parents[g][e] * N + col0 + s];
      x[(size_t)g * N + col0 + s] = v + r.normal();
    }
  }
}

}  // namespace

extern "C" int eddy_synth_config(int k, eddy_synth_cfg *c) {
  if (!c) return EDDY_ERR_ARG;
  std::memset(c, 0, sizeof(*c));
  c->mp = 3;
  c->perms = 100;
  c->prior_weight = 0.5;
  switch (k) {
    case 0:
      c->seed = 0xEDD10000ULL; c->genes = 200; c->nA = 100; c->nB = 100; c->rewire = 0.10;
      c->nsets = 1; c->gmin = 40; c->gmax = 40; c->prior_fixed = 20; c->prior_per_gene = 0;
      return EDDY_OK;
    case 1:
      c->seed = 0xEDD10001ULL; c->genes = 1000; c->nA = 200; c->nB = 200; c->rewire = 0.15;
      c->nsets = 1; c->gmin = 150; c->gmax = 150; c->prior_fixed = 60; c->prior_per_gene = 0;
      return EDDY_OK;
    case 2:
    case 3:
      c->seed = 0xEDD10002ULL; c->genes = 5000; c->nA = 300; c->nB = 300; c->rewire = 0.10;
      c->nsets = 500; c->gmin = 10; c->gmax = 300; c->prior_fixed = -1; c->prior_per_gene = 0.3;
      return EDDY_OK;
    case 4:
      c->seed = 0xEDD10004ULL; c->genes = 20000; c->nA = 300; c->nB = 900; c->rewire = 0.05;
      c->nsets = 2000; c->gmin = 5; c->gmax = 500; c->prior_fixed = -1; c->prior_per_gene = 0.3;
      return EDDY_OK;
  }
  return EDDY_ERR_ARG;
}

extern "C" int eddy_synth_dataset(uint64_t seed, int genes, int nA, int nB, double rewireFrac,
                                  uint8_t *X) {
  if (genes <= 0 || nA < 0 || nB < 0 || !X) return EDDY_ERR_ARG;
  int N = nA + nB;
  Rng r(seed);
  Dag a = make_dag(r, genes, 3);
  Dag b = rewire(r, a, rewireFrac);
  std::vector<double> x((size_t)genes * N);
  sample(r, a, nA, 0, N, x);
  sample(r, b, nB, nA, N, x);
  std::vector<int> idx(N);
  int cut1 = N / 3, cut2 = 2 * N / 3;
  for (int g = 0; g < genes; g++) {
    const double *row = &x[(size_t)g * N];
    for (int s = 0; s < N; s++) idx[s] = s;
    std::stable_sort(idx.begin(), idx.end(), [&](int p, int q) { return row[p] < row[q]; });
    for (int rk = 0; rk < N; rk++) X[(size_t)g * N + idx[rk]] = (uint8_t)(rk < cut1 ? 0 : rk < cut2 ? 1 : 2);
  }
  return EDDY_OK;
}

extern "C" int eddy_synth_sets(uint64_t seed, int genes, int nsets, int gmin, int gmax,
                               int prior_fixed, double prior_per_gene, int *set_sizes,
                               int *set_genes, long long genes_cap, int *prior_counts,
                               int *prior_pairs, long long pairs_cap) {
  if (genes <= 0 || nsets < 0 || gmin < 0 || gmax < gmin || gmax > genes) return EDDY_ERR_ARG;
  Rng r(seed ^ 0x5E75EED5ULL);
  std::vector<int> perm(genes);
  long long gpos = 0, ppos = 0;
  for (int s = 0; s < nsets; s++) {
    int G = gmin;
    if (gmax > gmin) {
      double lo = std::log((double)gmin), hi = std::log((double)gmax + 1.0);
      G = (int)std::floor(std::exp(lo + r.u01() * (hi - lo)));
      if (G < gmin) G = gmin;
      if (G > gmax) G = gmax;
    }
    if (gpos + G > genes_cap) return EDDY_ERR_CAPACITY;
    for (int i = 0; i < genes; i++) perm[i] = i;
    for (int i = 0; i < G; i++) {
      int j = i + (int)r.below(genes - i);
      std::swap(perm[i], perm[j]);
      set_genes[gpos + i] = perm[i];
    }
    long long npri = prior_fixed >= 0 ? prior_fixed : (long long)std::floor(prior_per_gene * G);
    long long maxp = (long long)G * (G - 1) / 2;
    if (npri > maxp) npri = maxp;
    if (ppos + npri > pairs_cap) return EDDY_ERR_CAPACITY;
    std::vector<uint8_t> used((size_t)G * G, 0);
    long long got = 0;
    for (long long tries = 0; got < npri && tries < 1000000; tries++) {
      int a = (int)r.below(G), b = (int)r.below(G);
      if (a == b || used[(size_t)a * G + b]) continue;
      used[(size_t)a * G + b] = used[(size_t)b * G + a] = 1;
      prior_pairs[2 * (ppos + got)] = a;
      prior_pairs[2 * (ppos + got) + 1] = b;
      got++;
    }
    set_sizes[s] = G;
    prior_counts[s] = (int)got;
    gpos += G;
    ppos += got;
  }
  return EDDY_OK;
}
