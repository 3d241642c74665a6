// textio.cpp - readers and writers of the reference's text formats.
//
// Quirks kept on purpose (they decide which rows and genes exist):
//   * the expression header's gene-id column is skipped and every later
//     tab-separated token is a sample, except that each token of length 1
//     lowers the sample count (main.cu:268-272, which is how a trailing tab's
//     "\n" token is dropped);
//   * the last expression row is dropped (numGenes--, main.cu:283);
//   * class 2 is the first label that differs from the first one, chosen
//     before the last label's newline is stripped (main.cu:312-327);
//   * gene-set genes missing from the data are dropped, the order of the rest
//     is kept, and the first data row with the same name is used
//     (main.cu:400-430);
//   * prior lines are read in 100-byte pieces, "neighbor-of" lines are
//     skipped, the last character of the third token is cut
//     (main.cu:433-503).
// Fixed-size limits of the reference (200 genes per set, 20500 genes, 50-char
// names) are lifted.
#include "textio.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <unordered_map>

#include <sys/stat.h>

namespace {

// strtok semantics on a private copy: runs of delimiters separate tokens,
// empty tokens are skipped
std::vector<std::string> split_tokens(const std::string &s, char delim) {
  std::vector<std::string> out;
  size_t i = 0, n = s.size();
  while (i < n) {
    while (i < n && s[i] == delim) i++;
    if (i >= n) break;
    size_t j = i;
    while (j < n && s[j] != delim) j++;
    out.push_back(s.substr(i, j - i));
    i = j;
  }
  return out;
}

bool read_line(FILE *f, std::string &line) {
  line.clear();
  int ch;
  bool any = false;
  while ((ch = fgetc(f)) != EOF) {
    any = true;
    line.push_back((char)ch);
    if (ch == '\n') break;
  }
  return any;
}

}  // namespace

// sscanf("%d") on the bytes [p, e) of one token, as the C library reads it:
// leading white space, an optional sign, at least one digit (the rest of the
// token is ignored); the number is read as a long (saturating like strtol)
// and stored into the int
static bool scan_int(const char *p, const char *e, int &v) {
  while (p < e && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t' || *p == '\v' || *p == '\f')) p++;
  bool neg = false;
  if (p < e && (*p == '+' || *p == '-')) neg = *p++ == '-';
  if (p >= e || *p < '0' || *p > '9') return false;
  unsigned long long x = 0;
  bool over = false;
  while (p < e && *p >= '0' && *p <= '9') {
    const unsigned d = (unsigned)(*p - '0');
    if (!over && x > (0x7FFFFFFFFFFFFFFFULL - d) / 10) over = true;
    if (!over) x = x * 10 + d;
    p++;
  }
  long long l;
  if (over) l = neg ? (-0x7FFFFFFFFFFFFFFFLL - 1) : 0x7FFFFFFFFFFFFFFFLL;
  else l = neg ? -(long long)x : (long long)x;
  v = (int)(uint32_t)(uint64_t)l;
  return true;
}

// The whole file in memory, scanned once: lines end at '\n' (kept in the
// line, as fgets-style reading keeps it), tokens are runs of non-tab bytes
// (strtok on '\t').  Same rows, values and errors as tokenising each line.
std::string load_expression(const std::string &path, ExprData &out) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return std::string("Expression file is NULL: ") + strerror(errno);
  std::vector<char> buf;
  {
    char tmp[1 << 16];
    size_t n;
    while ((n = fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
  }
  fclose(f);
  if (buf.empty()) return "Expression file is empty";
  const char *b0 = buf.data(), *be = b0 + buf.size();
  // line starts
  std::vector<size_t> ls;
  ls.push_back(0);
  for (size_t i = 0; i < buf.size(); i++)
    if (b0[i] == '\n' && i + 1 < buf.size()) ls.push_back(i + 1);
  const size_t nlines = ls.size();
  auto line_end = [&](size_t k) { return k + 1 < nlines ? b0 + ls[k + 1] : be; };
  // header: every token after the first is a sample, except 1-byte tokens
  int ns = 0, ntok = 0;
  for (const char *p = b0, *e = line_end(0); p < e;) {
    while (p < e && *p == '\t') p++;
    if (p >= e) break;
    const char *q = p;
    while (q < e && *q != '\t') q++;
    if (ntok++ > 0 && q - p != 1) ns++;
    p = q;
  }
  out.samples = ns;
  // data rows: lines 1 .. nlines-1, of which the last is dropped
  const int ng = nlines >= 2 ? (int)nlines - 2 : 0;
  out.genes.assign(ng, std::string());
  out.x.assign((size_t)ng * ns, 3);
  for (int g = 0; g < ng; g++) {
    const char *p = b0 + ls[g + 1], *e = line_end(g + 1);
    int k = 0;  // tokens seen
    for (const char *c = p; c < e && k <= ns;) {  // enough tokens? (checked before any value)
      while (c < e && *c == '\t') c++;
      if (c >= e) break;
      while (c < e && *c != '\t') c++;
      k++;
    }
    if (k < ns + 1) return "Expression row " + std::to_string(g + 2) + " has too few values";
    k = 0;
    uint8_t *row = out.x.data() + (size_t)g * ns;
    while (p < e && k <= ns) {
      while (p < e && *p == '\t') p++;
      if (p >= e) break;
      const char *q = p;
      while (q < e && *q != '\t') q++;
      if (k == 0) {
        out.genes[g].assign(p, q);
      } else {
        int v;
        if (!scan_int(p, q, v)) return "Expression row " + std::to_string(g + 2) + " has a non-integer value";
        row[k - 1] = (v >= -1 && v <= 1) ? (uint8_t)(v + 1) : 3;
      }
      k++;
      p = q;
    }
  }
  return "";
}

// ---- binary cache of a parsed expression file -----------------------------
// <input>.eddybin: "EDDYBIN1", the source's size and modification time, the
// sample and gene counts, the gene names (length-prefixed), the states.
static const char kCacheMagic[8] = {'E', 'D', 'D', 'Y', 'B', 'I', 'N', '1'};

static bool source_stamp(const std::string &path, long long &size, long long &mtime_ns) {
  struct stat st;
  if (stat(path.c_str(), &st) != 0) return false;
  size = (long long)st.st_size;
  mtime_ns = (long long)st.st_mtim.tv_sec * 1000000000LL + st.st_mtim.tv_nsec;
  return true;
}

bool load_expression_cache(const std::string &path, ExprData &out) {
  long long size, mt;
  if (!source_stamp(path, size, mt)) return false;
  FILE *f = fopen((path + ".eddybin").c_str(), "rb");
  if (!f) return false;
  char magic[8];
  long long hdr[4];
  bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, kCacheMagic, 8) == 0 && fread(hdr, sizeof(long long), 4, f) == 4 &&
            hdr[0] == size && hdr[1] == mt && hdr[2] >= 0 && hdr[3] >= 0;
  ExprData d;
  if (ok) {
    d.samples = (int)hdr[2];
    d.genes.resize((size_t)hdr[3]);
    for (auto &g : d.genes) {
      uint32_t n;
      if (fread(&n, 4, 1, f) != 1 || n > (1u << 20)) {
        ok = false;
        break;
      }
      g.resize(n);
      if (n && fread(&g[0], 1, n, f) != n) {
        ok = false;
        break;
      }
    }
  }
  if (ok) {
    d.x.resize((size_t)d.samples * d.genes.size());
    ok = d.x.empty() || fread(d.x.data(), 1, d.x.size(), f) == d.x.size();
  }
  fclose(f);
  if (ok) out = std::move(d);
  return ok;
}

void save_expression_cache(const std::string &path, const ExprData &d) {
  long long size, mt;
  if (!source_stamp(path, size, mt)) return;
  const std::string tmp = path + ".eddybin.tmp";
  FILE *f = fopen(tmp.c_str(), "wb");
  if (!f) return;
  long long hdr[4] = {size, mt, (long long)d.samples, (long long)d.genes.size()};
  bool ok = fwrite(kCacheMagic, 1, 8, f) == 8 && fwrite(hdr, sizeof(long long), 4, f) == 4;
  for (const auto &g : d.genes) {
    const uint32_t n = (uint32_t)g.size();
    ok = ok && fwrite(&n, 4, 1, f) == 1 && (n == 0 || fwrite(g.data(), 1, n, f) == n);
  }
  ok = ok && (d.x.empty() || fwrite(d.x.data(), 1, d.x.size(), f) == d.x.size());
  ok = fclose(f) == 0 && ok;
  if (ok) ok = rename(tmp.c_str(), (path + ".eddybin").c_str()) == 0;
  if (!ok) remove(tmp.c_str());
}

std::string load_classes(const std::string &path, int samples, ClassData &out) {
  FILE *f = fopen(path.c_str(), "r");
  if (!f) return std::string("Class File File is NULL: ") + strerror(errno);
  std::string line;
  bool ok = read_line(f, line);
  fclose(f);
  if (!ok) return "Class file is empty";
  std::vector<std::string> t = split_tokens(line, '\t');
  if ((int)t.size() < samples) return "Class file has fewer labels than samples";
  t.resize(samples);
  if (samples < 2) return "Class file needs two class labels";
  int i = 1;
  for (; i < samples; i++)
    if (t[i] != t[0]) break;
  if (i >= samples) return "Class file needs two class labels";
  out.class1 = t[0];
  out.class2 = t[i];
  std::string &last = t[samples - 1];
  if (!last.empty() && last.back() == '\n') last.pop_back();
  out.of.assign(samples, 2);
  out.n1 = out.n2 = 0;
  for (int s = 0; s < samples; s++) {
    if (t[s] == out.class1) {
      out.of[s] = 0;
      out.n1++;
    }
    if (t[s] == out.class2) {
      out.of[s] = 1;
      out.n2++;
    }
  }
  if (out.n1 + out.n2 != samples) return "Every sample must carry one of the two class labels";
  // e.g. "A\tA\tA\n": the last token "A\n" is picked as class 2 before its
  // newline is cut, leaving class 2 empty (main.cu:312-336)
  if (out.n1 == 0 || out.n2 == 0) return "Class file needs two class labels with samples each";
  return "";
}

std::string load_gene_sets(const std::string &path, const ExprData &expr, std::vector<GeneSet> &out) {
  FILE *f = fopen(path.c_str(), "r");
  if (!f) return std::string("Gene List File is NULL: ") + strerror(errno);
  std::unordered_map<std::string, int> first;
  for (int g = 0; g < (int)expr.genes.size(); g++) first.emplace(expr.genes[g], g);
  std::string line;
  while (read_line(f, line)) {
    if (!line.empty() && line[0] == '\n') continue;
    size_t nl = line.find('\n');
    std::string body = nl == std::string::npos ? line : line.substr(0, nl);
    std::vector<std::string> t = split_tokens(body, '\t');
    GeneSet s;
    s.name = t.empty() ? std::string() : t[0];
    s.listed = t.size() > 2 ? (int)t.size() - 2 : 0;
    for (size_t i = 2; i < t.size(); i++) {
      auto it = first.find(t[i]);
      if (it == first.end()) continue;
      s.rows.push_back(it->second);
      s.names.push_back(t[i]);
    }
    out.push_back(std::move(s));
  }
  fclose(f);
  return "";
}

void load_prior(const std::string &dir, GeneSet &set) {
  std::string path = dir + set.name + ".prior";
  FILE *f = fopen(path.c_str(), "r");
  set.prior_file = f != nullptr;
  if (!f) return;
  const int G = (int)set.names.size();
  std::unordered_map<std::string, int> pos;
  for (int k = G - 1; k >= 0; k--) pos[set.names[k]] = k;  // first match wins
  std::vector<uint8_t> seen((size_t)G * G, 0);
  char buf[100];
  while (fgets(buf, sizeof(buf), f)) {
    std::vector<std::string> t = split_tokens(std::string(buf), '\t');
    if (t.empty()) continue;
    auto a = pos.find(t[0]);
    if (a == pos.end()) continue;
    if (t.size() < 2) continue;
    if (t[1] == "neighbor-of") continue;
    if (t.size() < 3 || t[2].empty()) continue;
    std::string g2 = t[2].substr(0, t[2].size() - 1);
    auto b = pos.find(g2);
    if (b == pos.end()) continue;
    int r = a->second, c = b->second;
    if (!seen[(size_t)r * G + c]) {
      seen[(size_t)r * G + c] = seen[(size_t)c * G + r] = 1;
      set.prior.push_back(r);
      set.prior.push_back(c);
    }
  }
  fclose(f);
}

// ---- writers ---------------------------------------------------------------
void write_bdeu_scores(const std::string &path, const std::string &input, const std::string &classfile,
                       const std::string &setname, const std::string &class1, const std::string &class2, int U,
                       const std::vector<double> &lval) {
  FILE *f = fopen(path.c_str(), "w");
  if (!f) {
    printf("scoreFile is NULL. Error code : %s. Exiting...\n", strerror(errno));
    exit(EXIT_FAILURE);
  }
  fprintf(f, "\t\t BDEU SCORES\n");
  fprintf(f, "input : %s\n", input.c_str());
  fprintf(f, "class : %s\n", classfile.c_str());
  fprintf(f, "geneset : %s\n\n", setname.c_str());
  fprintf(f, "Network \t  %s  \t   \t    %s\n", class1.c_str(), class2.c_str());
  for (int j = 0; j < U; j++) fprintf(f, "Network %d\t %f \t %f \n", j + 1, lval[j], lval[j + U]);
  fclose(f);
  printf("Finished bdeu file %s\n", path.c_str());
}

void write_network_file(const std::string &path, const std::string &input, const std::string &classfile,
                        const std::string &setname, int U, const std::vector<int> &epn,
                        const std::vector<std::string> &genes, const std::vector<int> &nodes,
                        const std::vector<int> &edges, int total_edges, const std::vector<int> &ids) {
  FILE *f = fopen(path.c_str(), "w");
  if (!f) {
    printf("output when accessing networkOutput.txt is NULL. Error code : %s. Exiting...\n", strerror(errno));
    exit(EXIT_FAILURE);
  }
  const int G = (int)genes.size();
  fprintf(f, "%s", "\t\tUnique Networks\n\n");
  fprintf(f, "input : %s\n", input.c_str());
  fprintf(f, "class : %s\n", classfile.c_str());
  fprintf(f, "geneset : %s\n\n\n", setname.c_str());
  int pos = 0;
  for (int i = 0; i < U; i++) {
    int total = (i == U - 1) ? total_edges - epn[i] : epn[i + 1] - epn[i];
    fprintf(f, "network %d (%d)\t number of edges : %d\n", i + 1, ids[i], total);
    fprintf(f, "%s", "---------------");
    fprintf(f, "%s", "---------------\n\n");
    if (total == 0) continue;
    // parents of node j are nodes[j+1]-nodes[j]; the last node takes the rest
    for (int j = 0; j < G - 1; j++) {
      int change = nodes[(size_t)i * G + j + 1] - nodes[(size_t)i * G + j];
      if (change != 0) {
        total -= change;
        for (int k = 0; k < change; k++) fprintf(f, "%s - %s\n", genes[j].c_str(), genes[edges[pos++]].c_str());
      }
      if (j + 1 == G - 1 && total > 0)
        for (int k = 0; k < total; k++) fprintf(f, "%s - %s\n", genes[j + 1].c_str(), genes[edges[pos++]].c_str());
    }
    fprintf(f, "%s", "\n");
    fprintf(f, "%s", "\n");
  }
  fclose(f);
  printf("Finished writing %s\n", path.c_str());
}

void write_edge_list(const std::string &path, const std::vector<std::string> &genes, const std::vector<int> &nodes,
                     const std::vector<int> &edges, const int epn[3], const std::vector<int> &prior_pairs,
                     const std::string &class1, const std::string &class2) {
  FILE *f = fopen(path.c_str(), "w");
  if (!f) {
    printf("edge list file is NULL. Exiting...\n");
    exit(EXIT_FAILURE);
  }
  const int G = (int)genes.size();
  std::vector<uint8_t> prior((size_t)G * G, 0);
  for (size_t q = 0; q + 1 < prior_pairs.size(); q += 2) {
    prior[(size_t)prior_pairs[q] * G + prior_pairs[q + 1]] = 1;
    prior[(size_t)prior_pairs[q + 1] * G + prior_pairs[q]] = 1;
  }
  struct E {
    int node, edge, both;
  };
  std::vector<E> cl[2];
  int pos = 0;
  for (int i = 0; i < 2; i++) {
    int total = epn[i + 1] - epn[i];
    if (total == 0) continue;
    for (int j = 0; j < G - 1; j++) {
      int change = nodes[(size_t)i * G + j + 1] - nodes[(size_t)i * G + j];
      if (change != 0) {
        total -= change;
        for (int k = 0; k < change; k++) cl[i].push_back({j, edges[pos++], 0});
      }
      if (j + 1 == G - 1 && total > 0)
        for (int k = 0; k < total; k++) cl[i].push_back({j + 1, edges[pos++], 0});
    }
  }
  for (auto &a : cl[0])
    for (auto &b : cl[1])
      if (a.node == b.node && a.edge == b.edge) a.both = b.both = 1;
  static const char *ps[2] = {"NONE", "PRIOR"};
  std::vector<uint8_t> found((size_t)G * G, 0);
  for (auto &a : cl[0]) {
    fprintf(f, "%s\t%s\t%s\t%s\n", genes[a.node].c_str(), genes[a.edge].c_str(),
            a.both ? "Both" : class1.c_str(), ps[prior[(size_t)a.node * G + a.edge]]);
    found[(size_t)a.node * G + a.edge] = found[(size_t)a.edge * G + a.node] = 1;
  }
  for (auto &b : cl[1]) {
    if (b.both) continue;
    fprintf(f, "%s\t%s\t%s\t%s\n", genes[b.node].c_str(), genes[b.edge].c_str(), class2.c_str(),
            ps[prior[(size_t)b.node * G + b.edge]]);
    found[(size_t)b.node * G + b.edge] = found[(size_t)b.edge * G + b.node] = 1;
  }
  for (int r = 1; r < G; r++)
    for (int c = 0; c < r; c++)
      if (!found[(size_t)r * G + c] && prior[(size_t)r * G + c])
        fprintf(f, "%s\t%s\t%s\t%s\n", genes[r].c_str(), genes[c].c_str(), "Neither", "PRIOR");
  fclose(f);
  printf("Finished writing edgelist file %s\n", path.c_str());
}
