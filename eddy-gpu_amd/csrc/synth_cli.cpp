// eddy_synth - writes a synthetic workload in the reference's input formats:
//   <dir>/expr.txt    header "gene_id\tS0001...", one row per gene with
//                     values -1/0/1, plus one padding row (the loader drops
//                     the last row, main.cu:283)
//   <dir>/class.txt   one line of labels (classA x nA, classB x nB)
//   <dir>/sets.gmt    "SET_0001\thttp://synthetic/SET_0001\tG00012\t..."
//   <dir>/PRIORS/SET_0001.prior   "G00012\tinteracts-with\tG00345"
// usage: eddy_synth <workload 0..4> <dir> [nsets]
#include <sys/stat.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "eddy.h"

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: eddy_synth <workload 0..4> <dir> [nsets]\n");
    return 2;
  }
  eddy_synth_cfg c;
  if (eddy_synth_config(atoi(argv[1]), &c) != EDDY_OK) {
    fprintf(stderr, "unknown config\n");
    return 2;
  }
  std::string dir = argv[2];
  int nsets = argc > 3 ? atoi(argv[3]) : c.nsets;
  if (nsets > c.nsets || nsets < 0) nsets = c.nsets;
  mkdir(dir.c_str(), 0755);
  mkdir((dir + "/PRIORS").c_str(), 0755);
  const int N = c.nA + c.nB;
  std::vector<uint8_t> x((size_t)c.genes * N);
  if (eddy_synth_dataset(c.seed, c.genes, c.nA, c.nB, c.rewire, x.data()) != EDDY_OK) return 1;
  char name[32];
  FILE *f = fopen((dir + "/expr.txt").c_str(), "w");
  if (!f) return 1;
  fprintf(f, "gene_id");
  for (int s = 0; s < N; s++) fprintf(f, "\tS%04d", s + 1);
  fprintf(f, "\n");
  for (int g = 0; g <= c.genes; g++) {
    if (g < c.genes) snprintf(name, sizeof(name), "G%05d", g);
    fprintf(f, "%s", g < c.genes ? name : "PADDING");
    for (int s = 0; s < N; s++) fprintf(f, "\t%d", g < c.genes ? (int)x[(size_t)g * N + s] - 1 : 0);
    fprintf(f, "\n");
  }
  fclose(f);
  f = fopen((dir + "/class.txt").c_str(), "w");
  if (!f) return 1;
  for (int s = 0; s < N; s++) fprintf(f, "%s%s", s ? "\t" : "", s < c.nA ? "classA" : "classB");
  fprintf(f, "\n");
  fclose(f);
  long long gcap = (long long)c.nsets * c.gmax, pcap = (long long)c.nsets * c.gmax * 4 + 64;
  std::vector<int> sizes(c.nsets), genes(gcap), pc(c.nsets), pairs(2 * pcap);
  if (eddy_synth_sets(c.seed, c.genes, c.nsets, c.gmin, c.gmax, c.prior_fixed, c.prior_per_gene, sizes.data(),
                      genes.data(), gcap, pc.data(), pairs.data(), pcap) != EDDY_OK)
    return 1;
  f = fopen((dir + "/sets.gmt").c_str(), "w");
  if (!f) return 1;
  long long go = 0, po = 0;
  for (int s = 0; s < nsets; s++) {
    char sname[32];
    snprintf(sname, sizeof(sname), "SET_%04d", s + 1);
    fprintf(f, "%s\thttp://synthetic/%s", sname, sname);
    for (int i = 0; i < sizes[s]; i++) fprintf(f, "\tG%05d", genes[go + i]);
    fprintf(f, "\n");
    FILE *pf = fopen((dir + "/PRIORS/" + sname + ".prior").c_str(), "w");
    if (!pf) return 1;
    for (int q = 0; q < pc[s]; q++)
      fprintf(pf, "G%05d\tinteracts-with\tG%05d\n", genes[go + pairs[2 * (po + q)]], genes[go + pairs[2 * (po + q) + 1]]);
    fclose(pf);
    go += sizes[s];
    po += pc[s];
  }
  fclose(f);
  printf("wrote %s: %d genes x %d samples, %d sets\n", dir.c_str(), c.genes, N, nsets);
  return 0;
}
