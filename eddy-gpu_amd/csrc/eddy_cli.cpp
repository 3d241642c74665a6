// eddy - drop-in replacement of the reference `eddy-gpu` binary on one
// MI355X.  Same flags (main.cu:118-145), same inputs, same results.txt and
// per-set <set>_Networks.txt, <set>_BDEU_SCORES.txt, <set>_EdgeList.txt
// (main.cu:1504-1641, routines.cu:217-492).  All permutations of all gene sets
// run on the GPU through libeddy (include/eddy.h); this file only parses,
// calls the engine and writes files.
//
// Differences from the reference, all deliberate:
//   * `-p` (documented in the reference README but never parsed there) is an
//     alias of `-pD`;
//   * `--mode plain|batched`, `--tests all|cutoff` and `--no-learned` choose
//     the engine schedule (results are identical in every mode); by default
//     the first `--learn-sets K` (100) sets train the learned proposal's
//     network and the rest use it;
//     `--parse-only` prints the parsed inputs and stops (no GPU);
//   * batched runs pick their launch shapes with the autotuner first
//     (`--no-autotune` keeps the defaults; `--tune-cache FILE` keeps the
//     choice per device and dataset shape across runs);
//   * a degenerate beta fit (e.g. -r 1), where the reference's Boost call
//     throws and aborts, prints pDDN as nan and writes no files;
//   * inputs the reference would read out of bounds (missing class labels,
//     a third class, short expression rows) are rejected with EXIT_FAILURE.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "eddy.h"
#include "textio.h"

static void fail(const std::string &m) {
  printf("%s\n", m.c_str());
  exit(EXIT_FAILURE);
}

int main(int argc, char **argv) {
  std::string input, classfile, genesetfile;
  int parentCap = 0, perms = 100;
  double priorWeight = 0.5, alphaDDN = 0.05, theta = -1, lambda = 2.0, alphaEdge = 0.05;
  bool pAdjust = true;
  int mode = EDDY_MODE_BATCHED, schedule = EDDY_TESTS_CUTOFF, learned = 1, learn_sets = 100;
  bool parse_only = false, use_cache = false, autotune = true;
  std::string tune_cache;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    const char *nx = i + 1 < argc ? argv[i + 1] : "";
    if (a == "-d") input = nx;
    else if (a == "-g") genesetfile = nx;
    else if (a == "-c") classfile = nx;
    else if (a == "-l") lambda = atof(nx);
    else if (a == "-t") theta = atof(nx);
    else if (a == "-mp") parentCap = atoi(nx);
    else if (a == "-help" || a == "--help")
      printf("Required arguments : \n -d input.txt\n -c classfile.txt\n -g geneset.txt -mp # of parents\n"
             " -p pvalue for DDN significance testing\n");
    else if (a == "-r") perms = atoi(nx);
    else if (a == "-pw" || a == "pW") priorWeight = atof(nx);
    else if (a == "-pE") alphaEdge = atof(nx);
    else if (a == "-rawp") pAdjust = false;
    else if (a == "-pD" || a == "-p") alphaDDN = atof(nx);
    else if (a == "--mode") mode = std::string(nx) == "plain" ? EDDY_MODE_PLAIN : EDDY_MODE_BATCHED;
    else if (a == "--tests") schedule = std::string(nx) == "all" ? EDDY_TESTS_ALL : EDDY_TESTS_CUTOFF;
    else if (a == "--no-learned") learned = 0;
    else if (a == "--learn-sets") learn_sets = atoi(nx);
    else if (a == "--parse-only") parse_only = true;
    else if (a == "--cache") use_cache = true;
    else if (a == "--no-autotune") autotune = false;
    else if (a == "--tune-cache") tune_cache = nx;
  }
  if (input.empty()) fail("Invalid input file entered. Exiting...");
  if (classfile.empty()) fail("Invalid class file entered. Exiting...");
  if (genesetfile.empty()) fail("Invalid geneset file entered. Exiting...");
  printf("%s\n", genesetfile.c_str());
  if (alphaDDN < 0.0 || alphaDDN > 1.0) alphaDDN = .05;
  printf("alpha for signficance of DDN  = %f\n", alphaDDN);
  if (parentCap <= 0) parentCap = 3;
  printf("Max parents = %d\n", parentCap);
  if (perms <= 0) perms = 100;
  printf("Permutations = %d\n", perms);
  if (priorWeight < 0.0 || priorWeight > 1.0) priorWeight = 0.5;
  printf("priorWeight : %f\n", priorWeight);
  if (theta > 0)
    alphaEdge = 1 - pow(theta, 1.0 / lambda);
  else
    theta = pow(1 - alphaEdge, lambda);
  double alphaEdgePrior = 1 - (1 - alphaEdge) * pow(priorWeight, 1 / lambda);
  printf("alphaEdge : %f, alphaEdgePrior : %f\n", alphaEdge, alphaEdgePrior);
  printf("Multiple testing correction (edge detection) : %s\n", pAdjust ? "true" : "false");
  printf("[to be deprecated] lambda : %f, theta : %f\n", lambda, theta);

  ExprData expr;
  std::string e;
  if (use_cache && load_expression_cache(input, expr)) {
    printf("expression cache %s.eddybin\n", input.c_str());
  } else {
    e = load_expression(input, expr);
    if (!e.empty()) fail(e);
    if (use_cache) save_expression_cache(input, expr);
  }
  ClassData cls;
  e = load_classes(classfile, expr.samples, cls);
  if (!e.empty()) fail(e);
  std::vector<GeneSet> sets;
  e = load_gene_sets(genesetfile, expr, sets);
  if (!e.empty()) fail(e);
  FILE *res = fopen("results.txt", "w");
  if (!res) fail(std::string("results file is NULL: ") + strerror(errno));
  printf("%d samples\n%d genes\n", expr.samples, (int)expr.genes.size());
  printf("Classes: %d %s, %d %s\n", cls.n1, cls.class1.c_str(), cls.n2, cls.class2.c_str());
  for (auto &s : sets) load_prior("PRIORS/", s);
  if (parse_only) {  // inputs as the engine would receive them; no GPU needed
    uint64_t h = 1469598103934665603ULL;  // FNV-1a of the states and the gene names
    for (uint8_t v : expr.x) h = (h ^ v) * 1099511628211ULL;
    for (const auto &gname : expr.genes) {
      for (char ch : gname) h = (h ^ (uint8_t)ch) * 1099511628211ULL;
      h = (h ^ 0xFFu) * 1099511628211ULL;
    }
    printf("expression hash %016llx\n", (unsigned long long)h);
    for (auto &s : sets) {
      printf("set %s genes %d prior %d :", s.name.c_str(), (int)s.rows.size(), (int)s.prior.size() / 2);
      for (int r : s.rows) printf(" %d", r);
      printf("\n");
    }
    fclose(res);
    return 0;
  }

  eddy_ctx *ctx = nullptr;
  int rc = eddy_ctx_create(0, &ctx);
  if (rc != EDDY_OK) fail("no usable GPU (eddy_ctx_create " + std::to_string(rc) + ")");
  rc = eddy_upload_dataset(ctx, expr.x.data(), (int)expr.genes.size(), expr.samples, cls.of.data());
  if (rc != EDDY_OK) fail(std::string("dataset upload failed: ") + eddy_last_error(ctx));
  eddy_params p{};
  p.mp = parentCap;
  p.perms = perms;
  p.p_adjust = pAdjust ? 1 : 0;
  p.alpha_edge = alphaEdge;
  p.alpha_edge_prior = alphaEdgePrior;
  p.alpha_ddn = alphaDDN;
  p.mode = mode;
  p.schedule = schedule;
  p.learned = 0;
  std::vector<eddy_set_desc> desc(sets.size());
  for (size_t i = 0; i < sets.size(); i++) {
    desc[i].n_genes = (int)sets[i].rows.size();
    desc[i].genes = sets[i].rows.data();
    desc[i].n_prior = (int)sets[i].prior.size() / 2;
    desc[i].prior = sets[i].prior.data();
  }
  std::vector<eddy_set_result> out(sets.size());
  if (autotune && mode == EDDY_MODE_BATCHED && !desc.empty()) {
    eddy_launch_cfg lcfg{};
    eddy_tune_report trep{};
    rc = eddy_autotune(ctx, desc.data(), (int)desc.size(), &p, tune_cache.empty() ? nullptr : tune_cache.c_str(), &lcfg,
                       &trep);
    if (rc != EDDY_OK) fail(std::string("autotune failed: ") + eddy_last_error(ctx));
    fprintf(stderr, "launch shapes (%s, %.0f ms): graph %d tile_order %d item_order %d loo_waves %d bdeu_groups %d "
            "bdeu_blocks %d\n", trep.from_cache ? "cached" : "tuned", trep.tune_ms, lcfg.graph, lcfg.tile_order,
            lcfg.item_order, lcfg.loo_waves, lcfg.bdeu_groups, lcfg.bdeu_blocks);
  }
  if (learned && mode == EDDY_MODE_BATCHED && schedule == EDDY_TESTS_CUTOFF && learn_sets >= 0) {
    eddy_learn_cfg lc{};
    lc.train_sets = learn_sets;
    lc.epochs = 1;
    lc.batch = 16384;
    lc.max_rows = 65536;
    lc.label_perms = 10;
    lc.lr = 3e-3f;
    lc.margin = 3.0f;
    lc.seed = 1;
    rc = eddy_run_learned(ctx, desc.data(), (int)desc.size(), &p, &lc, out.data(), nullptr);
  } else {
    rc = eddy_run_sets(ctx, desc.data(), (int)desc.size(), &p, out.data());
  }
  if (rc != EDDY_OK) fail(std::string("engine failed: ") + eddy_last_error(ctx));

  fprintf(res, "Pathway\tJS\tPval\tnGenes\n");
  for (size_t i = 0; i < sets.size(); i++) {
    const GeneSet &s = sets[i];
    const eddy_set_result &r = out[i];
    const int G = (int)s.rows.size();
    // the reference's per-set lines in its order (main.cu:390-537, 1124, 1432, 1462)
    printf("%s\n%d genes in geneset\n", s.name.c_str(), s.listed);
    printf("Adjusted genes : %d\n", G);
    const std::string prior_path = "PRIORS/" + s.name + ".prior";
    if (prior_path.find('\r') != std::string::npos) printf("File problem! Uses window endings!\n");
    printf("file : %s\n", prior_path.c_str());
    if (!s.prior_file) printf("No prior file exists. Computing without prior knowledge\n");
    printf("Permutations begin:\n");
    if (r.U0 > 0) printf("Original Number of unique Networks : %d\n", r.U0);
    if (r.status == EDDY_SET_FAILED_NAN && r.perms_run == 1) printf("jsVal[0] NAN--Breaking permutation loop\n");
    printf("permutation loop (%d) finished\n", perms);
    if (r.status == EDDY_SET_FAILED_NAN) {
      fprintf(res, "%s failed with %d genes- one of the js scores was nan\n", s.name.c_str(), G);
      printf("nan value- quiting!\n\n\n");
      continue;
    }
    printf("Original JS : %f\n", r.js0);
    printf("alpha : %f beta : %f\n", r.alpha, r.beta);
    if (r.status == EDDY_SET_DOMAIN)
      printf("ibetac domain error (alpha=%g beta=%g): pDDN reported as nan\n", r.alpha, r.beta);
    printf("pDDN = %f\n", r.pDDN);
    fprintf(res, "%s\t%f\t%f\t%d\n", s.name.c_str(), r.js0, r.pDDN, G);
    if (!(r.pDDN < alphaDDN)) {
      printf("\nPathway finished.\n\n");
      continue;
    }
    const int U = r.U0;
    std::vector<int> ids(U), nodes((size_t)U * G), epn(U + 1), edges((size_t)U * G * std::min(parentCap, G) + 1);
    std::vector<double> lval(2 * U);
    int ne = 0;
    rc = eddy_get_perm0_networks(ctx, (int)i, ids.data(), nodes.data(), (int)nodes.size(), epn.data(), edges.data(),
                                 (int)edges.size(), lval.data(), &ne);
    if (rc != EDDY_OK) fail(std::string("engine failed: ") + eddy_last_error(ctx));
    write_network_file(s.name + "_Networks.txt", input, classfile, s.name, U, epn, s.names, nodes, edges, ne, ids);
    write_bdeu_scores(s.name + "_BDEU_SCORES.txt", input, classfile, s.name, cls.class1, cls.class2, U, lval);
    std::vector<int> en(2 * G + 1), ee((size_t)G * G + 1);
    int eepn[3];
    rc = eddy_get_edgelist(ctx, (int)i, en.data(), eepn, ee.data(), (int)ee.size());
    if (rc != EDDY_OK) fail(std::string("engine failed: ") + eddy_last_error(ctx));
    write_edge_list(s.name + "_EdgeList.txt", s.names, en, ee, eepn, s.prior, cls.class1, cls.class2);
    printf("\nPathway finished.\n\n");
  }
  fclose(res);
  eddy_timing t{};
  eddy_get_timing(ctx, &t);
  printf("Total run time : %f\n", t.total_ms);
  eddy_ctx_destroy(ctx);
  return 0;
}
