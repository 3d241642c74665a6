// textio.h - the reference's text formats: expression matrix, class file, GMT
// gene sets, PRIORS/<set>.prior (main.cu:206-503) and the three per-set output
// files plus results.txt (routines.cu:217-492, main.cu:358, 1465-1502).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

struct ExprData {
  std::vector<std::string> genes;   // row names (first column)
  int samples = 0;                  // after the 1-character-token rule
  std::vector<uint8_t> x;           // [genes x samples] states 0,1,2 (value+1), 3 otherwise
};

struct ClassData {
  std::string class1, class2;
  std::vector<uint8_t> of;          // [samples] 0 = class1, 1 = class2
  int n1 = 0, n2 = 0;
};

struct GeneSet {
  std::string name;
  int listed = 0;                   // genes named on the GMT line (before the lookup)
  std::vector<int> rows;            // dataset rows, set order (missing genes dropped)
  std::vector<std::string> names;   // names of the kept genes
  std::vector<int> prior;           // pairs (a, b) of set positions
  bool prior_file = false;          // PRIORS/<name>.prior was found
};

// Each loader returns an empty string on success, else the reference's error
// message (the CLI prints it and exits with EXIT_FAILURE).
std::string load_expression(const std::string &path, ExprData &out);
// Binary cache of a parsed expression file at <path>.eddybin, valid while the
// source's size and modification time are unchanged (no counterpart in the
// reference; the CLI uses it only with --cache).
bool load_expression_cache(const std::string &path, ExprData &out);
void save_expression_cache(const std::string &path, const ExprData &d);
std::string load_classes(const std::string &path, int samples, ClassData &out);
std::string load_gene_sets(const std::string &path, const ExprData &expr, std::vector<GeneSet> &out);
void load_prior(const std::string &dir, GeneSet &set);

void write_bdeu_scores(const std::string &path, const std::string &input, const std::string &classfile,
                       const std::string &setname, const std::string &class1, const std::string &class2, int U,
                       const std::vector<double> &lval);
void write_network_file(const std::string &path, const std::string &input, const std::string &classfile,
                        const std::string &setname, int U, const std::vector<int> &epn,
                        const std::vector<std::string> &genes, const std::vector<int> &nodes,
                        const std::vector<int> &edges, int total_edges, const std::vector<int> &ids);
void write_edge_list(const std::string &path, const std::vector<std::string> &genes, const std::vector<int> &nodes,
                     const std::vector<int> &edges, const int epn[3], const std::vector<int> &prior_pairs,
                     const std::string &class1, const std::string &class2);
