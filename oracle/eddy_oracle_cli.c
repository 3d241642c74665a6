/*
 * eddy_oracle_cli.c - the oracle's restatement of the reference `eddy-gpu`
 * process: argv parsing (main.cu:58-204), input loaders (main.cu:206-503),
 * the gene-set loop (main.cu:359-1700) and the three output writers
 * (routines.cu:217-492).  TEST INFRASTRUCTURE ONLY: the golden-file tests run
 * it to produce the expected results.txt / _Networks / _BDEU_SCORES /
 * _EdgeList files that the product's `eddy` CLI must reproduce byte for byte.
 *
 * Fixed-size limits of the reference (MAX_GENESET_SIZE 200, MAX_NUM_GENES,
 * MAX_NUM_SAMPLES, 50-char names, definitions.cuh:17-22) are lifted; all
 * parsing quirks are kept (last expression row dropped, 1-character header
 * tokens dropped, 100-byte prior lines, `-p` not parsed).  Where the reference
 * has undefined behaviour (missing tokens, a third class label) this prints an
 * error and exits with EXIT_FAILURE.  Where Boost would throw a domain error
 * (degenerate beta fit) the row is written with pDDN = nan and no files.
 */
#define _GNU_SOURCE
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "eddy_oracle.h"

static void die(const char *msg) {
  printf("%s\n", msg);
  exit(EXIT_FAILURE);
}

static char *xstrdup(const char *s) {
  char *d = strdup(s ? s : "");
  if (!d) die("out of memory");
  return d;
}

/* read one line of any length (the reference's fgets with a 1.5 MB buffer) */
static char *read_line(FILE *f) {
  char *line = NULL;
  size_t cap = 0;
  ssize_t n = getline(&line, &cap, f);
  if (n < 0) {
    free(line);
    return NULL;
  }
  return line;
}

/* ---- writers: routines.cu:217-238 --------------------------------------- */
static void write_bdeu(const char *path, const char *inputFile,
                       const char *classFile, const char *setName,
                       const char *class1, const char *class2, int U,
                       const double *lval) {
  FILE *f = fopen(path, "w");
  if (!f) die("scoreFile is NULL. Exiting...");
  fprintf(f, "\t\t BDEU SCORES\n");
  fprintf(f, "input : %s\n", inputFile);
  fprintf(f, "class : %s\n", classFile);
  fprintf(f, "geneset : %s\n\n", setName);
  fprintf(f, "Network \t  %s  \t   \t    %s\n", class1, class2);
  for (int j = 0; j < U; j++)
    fprintf(f, "Network %d\t %f \t %f \n", j + 1, lval[j], lval[j + U]);
  fclose(f);
  printf("Finished bdeu file %s\n", path);
}

/* routines.cu:240-306 */
static void write_networks(const char *path, const char *inputFile,
                           const char *classFile, const char *setName, int U,
                           const int *epn, char **genes, int G,
                           const int *nodes, const int *edges, int totalEdges,
                           const int *ids) {
  FILE *f = fopen(path, "w");
  if (!f) die("output when accessing networkOutput.txt is NULL. Exiting...");
  fprintf(f, "%s", "\t\tUnique Networks\n\n");
  fprintf(f, "input : %s\n", inputFile);
  fprintf(f, "class : %s\n", classFile);
  fprintf(f, "geneset : %s\n\n\n", setName);
  int edgePos = 0;
  for (int i = 0; i < U; i++) {
    int total = (i == U - 1) ? totalEdges - epn[i] : epn[i + 1] - epn[i];
    fprintf(f, "network %d (%d)\t number of edges : %d\n", i + 1, ids[i], total);
    fprintf(f, "%s", "---------------");
    fprintf(f, "%s", "---------------\n\n");
    if (total == 0) continue;
    for (int j = 0; j < G - 1; j++) {
      if (nodes[i * G + j + 1] != nodes[i * G + j]) {
        int change = nodes[i * G + j + 1] - nodes[i * G + j];
        total -= change;
        for (int k = 0; k < change; k++)
          fprintf(f, "%s - %s\n", genes[j], genes[edges[edgePos++]]);
      }
      if (j + 1 == G - 1 && total > 0) {
        for (int k = 0; k < total; k++)
          fprintf(f, "%s - %s\n", genes[j + 1], genes[edges[edgePos++]]);
      }
    }
    fprintf(f, "%s", "\n");
    fprintf(f, "%s", "\n");
  }
  fclose(f);
  printf("Finished writing %s\n", path);
}

/* routines.cu:308-492 */
typedef struct {
  int node, edge, inBoth;
} gnode;

static void write_edgelist(const char *path, char **genes, int G,
                           const int *nodes, const int *edges,
                           const int *edgesPN, const int *prior,
                           const char *class1, const char *class2) {
  FILE *f = fopen(path, "w");
  if (!f) die("edge list file is NULL. Exiting...");
  int cap = edgesPN[2] + 1;
  gnode *c1 = (gnode *)malloc(sizeof(gnode) * cap);
  gnode *c2 = (gnode *)malloc(sizeof(gnode) * cap);
  int n1 = 0, n2 = 0, edgePos = 0;
  for (int i = 0; i < 2; i++) {
    int total = edgesPN[i + 1] - edgesPN[i];
    if (total == 0) continue;
    for (int j = 0; j < G - 1; j++) {
      if (nodes[i * G + j + 1] != nodes[i * G + j]) {
        int change = nodes[i * G + j + 1] - nodes[i * G + j];
        total -= change;
        for (int k = 0; k < change; k++) {
          gnode t = {j, edges[edgePos++], 0};
          if (i == 0) c1[n1++] = t; else c2[n2++] = t;
        }
      }
      if (j + 1 == G - 1 && total > 0) {
        for (int k = 0; k < total; k++) {
          gnode t = {j + 1, edges[edgePos++], 0};
          if (i == 0) c1[n1++] = t; else c2[n2++] = t;
        }
      }
    }
  }
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n2; j++)
      if (c1[i].node == c2[j].node && c1[i].edge == c2[j].edge) {
        c1[i].inBoth = 1;
        c2[j].inBoth = 1;
      }
  static const char *priorString[2] = {"NONE", "PRIOR"};
  char *found = (char *)calloc((size_t)G * G + 1, 1);
  for (int i = 0; i < n1; i++) {
    int row = c1[i].node, col = c1[i].edge;
    fprintf(f, "%s\t%s\t%s\t%s\n", genes[row], genes[col],
            c1[i].inBoth == 1 ? "Both" : class1, priorString[prior[row * G + col]]);
    found[row * G + col] = found[col * G + row] = 1;
  }
  for (int i = 0; i < n2; i++) {
    int row = c2[i].node, col = c2[i].edge;
    if (c2[i].inBoth != 1) {
      fprintf(f, "%s\t%s\t%s\t%s\n", genes[row], genes[col], class2,
              priorString[prior[row * G + col]]);
      found[row * G + col] = found[col * G + row] = 1;
    }
  }
  for (int row = 1; row < G; row++)
    for (int col = 0; col < row; col++)
      if (!found[row * G + col] && prior[row * G + col] == 1)
        fprintf(f, "%s\t%s\t%s\t%s\n", genes[row], genes[col], "Neither", "PRIOR");
  free(found);
  free(c1);
  free(c2);
  fclose(f);
  printf("Finished writing edgelist file %s\n", path);
}

/* uncapped CSR of perm 0's two full networks: main.cu:1530-1612 */
static void uncapped_pair(const uint8_t *bits0, const uint8_t *bitsL, int G,
                          int *nodes, int **edges_out, int edgesPN[3]) {
  int c = G * (G - 1) / 2;
  const uint8_t *b[2] = {bits0, bitsL};
  int tot[2] = {0, 0};
  for (int i = 0; i < 2; i++)
    for (int p = 0; p < c; p++) tot[i] += b[i][p];
  edgesPN[0] = 0;
  edgesPN[1] = tot[0];
  edgesPN[2] = tot[0] + tot[1];
  int *edges = (int *)malloc(sizeof(int) * (edgesPN[2] + 1));
  int pos = 0;
  for (int i = 0; i < 2; i++) {
    int acc = 0, pr = 0;
    for (int t = 0; t < G; t++) {
      nodes[i * G + t] = acc;
      for (int col = 0; col < t; col++, pr++)
        if (b[i][pr]) {
          edges[pos++] = col;
          acc++;
        }
    }
  }
  *edges_out = edges;
}

int eo_uncapped_pair(const uint8_t *bits0, const uint8_t *bitsL, int G, int *nodes, int *edges, int edges_cap,
                     int edgesPN[3]) {
  int *e = NULL;
  uncapped_pair(bits0, bitsL, G, nodes, &e, edgesPN);
  int n = edgesPN[2];
  if (n > edges_cap) {
    free(e);
    return -1;
  }
  memcpy(edges, e, sizeof(int) * (size_t)n);
  free(e);
  return n;
}

int eo_main(int argc, char **argv) {
  char *inputFile = NULL, *classFile = NULL, *genesetFile = NULL;
  int parentCap = 0, perms = 100, flag_pAdjust = 1;
  double priorWeight = 0.5, alphaDDN = 0.05, theta = -1, lambda = 2.0;
  double alphaEdge = 0.05;
  /* main.cu:118-145 */
  for (int i = 1; i < argc; i++) {
    const char *nx = i + 1 < argc ? argv[i + 1] : NULL;
    if (strcmp(argv[i], "-d") == 0) inputFile = (char *)nx;
    else if (strcmp(argv[i], "-g") == 0) genesetFile = (char *)nx;
    else if (strcmp(argv[i], "-c") == 0) classFile = (char *)nx;
    else if (strcmp(argv[i], "-l") == 0) lambda = nx ? atof(nx) : 0;
    else if (strcmp(argv[i], "-t") == 0) theta = nx ? atof(nx) : 0;
    else if (strcmp(argv[i], "-mp") == 0) parentCap = nx ? atoi(nx) : 0;
    else if (strcmp(argv[i], "-help") == 0 || strcmp(argv[i], "--help") == 0)
      printf("Required arguments : \n -d input.txt\n -c classfile.txt\n -g "
             "geneset.txt -mp # of parents\n -p pvalue for DDN significance "
             "testing\n");
    else if (strcmp(argv[i], "-r") == 0) perms = nx ? atoi(nx) : 0;
    else if (strcmp(argv[i], "-pw") == 0 || strcmp(argv[i], "pW") == 0) priorWeight = nx ? atof(nx) : 0;
    else if (strcmp(argv[i], "-pE") == 0) alphaEdge = nx ? atof(nx) : 0;
    else if (strcmp(argv[i], "-rawp") == 0) flag_pAdjust = 0;
    else if (strcmp(argv[i], "-pD") == 0) alphaDDN = nx ? atof(nx) : 0;
  }
  if (!inputFile) die("Invalid input file entered. Exiting...");
  if (!classFile) die("Invalid class file entered. Exiting...");
  if (!genesetFile) die("Invalid geneset file entered. Exiting...");
  if (alphaDDN < 0.0 || alphaDDN > 1.0) alphaDDN = .05;
  if (parentCap <= 0) parentCap = 3;
  if (perms <= 0) perms = 100;
  if (priorWeight < 0.0 || priorWeight > 1.0) priorWeight = 0.5;
  if (theta > 0) alphaEdge = 1 - pow(theta, 1.0 / lambda);
  double alphaEdgePrior = 1 - (1 - alphaEdge) * pow(priorWeight, 1 / lambda);

  FILE *fpExpr = fopen(inputFile, "r");
  FILE *fpClass = fopen(classFile, "r");
  FILE *fpGeneSet = fopen(genesetFile, "r");
  FILE *fpResults = fopen("results.txt", "w");
  if (!fpExpr) die("Expression file is NULL");
  if (!fpClass) die("Class File File is NULL");
  if (!fpGeneSet) die("Gene List File is NULL");
  if (!fpResults) die("results file is NULL");

  /* expression: main.cu:255-301 */
  char *line = read_line(fpExpr);
  if (!line) die("empty expression file");
  int capS = 16, numSamples = 0;
  char **sampleNames = (char **)malloc(sizeof(char *) * capS);
  char *tok = strtok(line, "\t");
  tok = strtok(NULL, "\t");
  while (tok) {
    if (numSamples == capS) sampleNames = (char **)realloc(sampleNames, sizeof(char *) * (capS *= 2));
    sampleNames[numSamples++] = xstrdup(tok);
    tok = strtok(NULL, "\t");
  }
  free(line);
  for (int i = 0; i < numSamples; i++)
    if (strlen(sampleNames[i]) == 1) numSamples--;
  int capG = 1024, numGenes = 0;
  char **geneNames = (char **)malloc(sizeof(char *) * capG);
  int *data = NULL;
  size_t dcap = 0;
  while ((line = read_line(fpExpr))) {
    if (numGenes == capG) geneNames = (char **)realloc(geneNames, sizeof(char *) * (capG *= 2));
    tok = strtok(line, "\t");
    geneNames[numGenes] = xstrdup(tok);
    if ((size_t)(numGenes + 1) * numSamples > dcap) {
      dcap = ((size_t)(numGenes + 1) * numSamples) * 2 + 16;
      data = (int *)realloc(data, sizeof(int) * dcap);
    }
    int ok = 1;
    for (int j = 0; j < numSamples; j++) {
      tok = strtok(NULL, "\t");
      if (!tok || sscanf(tok, "%d", &data[(size_t)numGenes * numSamples + j]) != 1) {
        ok = 0;
        break;
      }
    }
    numGenes++;
    free(line);
    if (!ok) {
      /* a short row is only tolerated as the (dropped) last row */
      line = read_line(fpExpr);
      if (line) die("malformed expression row");
      break;
    }
  }
  numGenes--;
  fclose(fpExpr);
  if (numGenes < 0) numGenes = 0;

  /* class file: main.cu:303-352 */
  line = read_line(fpClass);
  if (!line) die("empty class file");
  char **classNames = (char **)malloc(sizeof(char *) * (numSamples + 1));
  tok = strtok(line, "\t");
  for (int i = 0; i < numSamples; i++) {
    if (!tok) die("class file has fewer labels than samples");
    classNames[i] = xstrdup(tok);
    tok = strtok(NULL, "\t");
  }
  free(line);
  fclose(fpClass);
  int i2;
  for (i2 = 1; i2 < numSamples; i2++)
    if (strcmp(classNames[i2], classNames[0])) break;
  if (numSamples < 2 || i2 >= numSamples) die("class file needs two class labels");
  char *class1 = xstrdup(classNames[0]);
  char *class2 = xstrdup(classNames[i2]);
  {
    char *last = classNames[numSamples - 1];
    size_t L = strlen(last);
    if (L && last[L - 1] == '\n') last[L - 1] = '\0';
  }
  int numClass1 = 0, numClass2 = 0;
  int *classIDs = (int *)malloc(sizeof(int) * numSamples);
  for (int i = 0; i < numSamples; i++) {
    classIDs[i] = -1;
    if (!strcmp(classNames[i], class1)) { numClass1++; classIDs[i] = 0; }
    if (!strcmp(classNames[i], class2)) { numClass2++; classIDs[i] = 1; }
  }
  if (numClass1 + numClass2 != numSamples) die("every sample must carry one of the two class labels");
  printf("Classes: %d %s, %d %s\n", numClass1, class1, numClass2, class2);

  eo_params prm;
  prm.mp = parentCap;
  prm.perms = perms;
  prm.pAdjust = flag_pAdjust;
  prm.alphaEdge = alphaEdge;
  prm.alphaEdgePrior = alphaEdgePrior;
  prm.alphaDDN = alphaDDN;
  prm.threads = 1;
  {
    const char *th = getenv("EDDY_ORACLE_THREADS");
    if (th) prm.threads = atoi(th);
  }
  eo_reset_stream();

  fprintf(fpResults, "Pathway\tJS\tPval\tnGenes\n");
  while ((line = read_line(fpGeneSet))) {
    if (line[0] == '\n') { free(line); continue; }
    /* main.cu:375-393 */
    char *pathwayName = strtok(line, "\n");
    tok = strtok(pathwayName, "\t");
    tok = strtok(NULL, "\t");
    tok = strtok(NULL, "\t");
    int glen = 0, gcap = 16;
    char **gsGenes = (char **)malloc(sizeof(char *) * gcap);
    while (tok) {
      if (glen == gcap) gsGenes = (char **)realloc(gsGenes, sizeof(char *) * (gcap *= 2));
      gsGenes[glen++] = tok;
      tok = strtok(NULL, "\t");
    }
    printf("%s\n", pathwayName);
    /* main.cu:400-430 */
    int *setIdx = (int *)malloc(sizeof(int) * (glen + 1));
    char **genes = (char **)malloc(sizeof(char *) * (glen + 1));
    int G = 0;
    for (int i = 0; i < glen; i++)
      for (int j = 0; j < numGenes; j++)
        if (!strcmp(geneNames[j], gsGenes[i])) {
          genes[G] = geneNames[j];
          setIdx[G++] = j;
          break;
        }
    /* prior: main.cu:433-503 */
    int *prior = (int *)calloc((size_t)(G > 0 ? G : 1) * (G > 0 ? G : 1), sizeof(int));
    {
      char path[4096];
      snprintf(path, sizeof(path), "PRIORS/%s.prior", pathwayName);
      FILE *pf = fopen(path, "r");
      if (!pf) printf("No prior file exists. Computing without prior knowledge\n");
      char pb[100];
      while (pf && fgets(pb, sizeof(pb), pf)) {
        char *t1 = strtok(pb, "\t");
        int row = -1, col = -1;
        for (int k = 0; t1 && k < G; k++)
          if (strcmp(genes[k], t1) == 0) { row = k; break; }
        if (row < 0) continue;
        char *t2 = strtok(NULL, "\t");
        if (!t2) continue;
        if (strcmp(t2, "neighbor-of") == 0) continue;
        char *t3 = strtok(NULL, "\t");
        if (!t3 || !*t3) continue;
        t3[strlen(t3) - 1] = '\0';
        for (int k = 0; k < G; k++)
          if (strcmp(genes[k], t3) == 0) { col = k; break; }
        if (col < 0) continue;
        prior[row * G + col] = 1;
        prior[col * G + row] = 1;
      }
      if (pf) fclose(pf);
    }

    /* permutation loop: main.cu:538-1460 */
    int S = numSamples + 2, c = G * (G - 1) / 2;
    double *js = (double *)calloc(perms, sizeof(double));
    int *cls = (int *)malloc(sizeof(int) * numSamples);
    memcpy(cls, classIDs, sizeof(int) * numSamples);
    int *t1 = (int *)malloc(sizeof(int) * (size_t)(G ? G : 1) * numClass1);
    int *t2 = (int *)malloc(sizeof(int) * (size_t)(G ? G : 1) * numClass2);
    /* perm-0 captures */
    size_t ebound = (size_t)S * (G ? G : 1) * (parentCap < G ? parentCap : (G ? G : 1)) + 1;
    uint8_t *bits0 = (uint8_t *)malloc((size_t)S * (c ? c : 1));
    int *f_ids = (int *)malloc(sizeof(int) * S);
    int *f_nodes = (int *)malloc(sizeof(int) * (size_t)S * (G ? G : 1));
    int *f_edges = (int *)malloc(sizeof(int) * ebound);
    int *f_epn = (int *)malloc(sizeof(int) * (S + 1));
    double *f_lval = (double *)malloc(sizeof(double) * 2 * S);
    int f_U = 0, f_edSum = 0;
    for (int perm = 0; perm < perms; perm++) {
      eo_perm_labels(numSamples, numClass1, perm, cls);
      eo_class_split(data, numSamples, setIdx, G, cls, numClass1, numClass2, t1, t2);
      eo_perm_out o;
      memset(&o, 0, sizeof(o));
      if (perm == 0) {
        o.edgebits = bits0;
        o.uni_ids = f_ids;
        o.uni_nodes = f_nodes;
        o.uni_edges = f_edges;
        o.uni_epn = f_epn;
        o.lval = f_lval;
      }
      eo_run_perm(t1, t2, G, numClass1, numClass2, prior, &prm, &o);
      if (perm == 0) {
        f_U = o.U;
        f_edSum = o.edSum;
        printf("Original Number of unique Networks : %d\n", o.U);
      }
      js[perm] = o.js;
      if (isnan(js[0])) {
        printf("jsVal[0] NAN--Breaking permutation loop\n");
        js[0] = -999.0;
        break;
      }
    }
    int nanFlag = 0;
    for (int i = 0; i < perms; i++)
      if (isnan(js[i])) nanFlag = 1;
    if (js[0] < 0 || nanFlag) {
      fprintf(fpResults, "%s failed with %d genes- one of the js scores was nan\n", pathwayName, G);
    } else {
      printf("Original JS : %f\n", js[0]);
      double a, b;
      int de;
      double pDDN = eo_pvalue(js, perms, &a, &b, &de);
      printf("alpha : %f beta : %f\n", a, b);
      if (de) printf("ibetac domain error (alpha=%g beta=%g): pDDN reported as nan\n", a, b);
      printf("pDDN = %f\n", pDDN);
      fprintf(fpResults, "%s\t%f\t%f\t%d\n", pathwayName, js[0], pDDN, G);
      if (pDDN < alphaDDN) {
        char path[4200];
        snprintf(path, sizeof(path), "%s_Networks.txt", pathwayName);
        write_networks(path, inputFile, classFile, pathwayName, f_U, f_epn, genes, G,
                       f_nodes, f_edges, f_edSum, f_ids);
        snprintf(path, sizeof(path), "%s_BDEU_SCORES.txt", pathwayName);
        write_bdeu(path, inputFile, classFile, pathwayName, class1, class2, f_U, f_lval);
        int *nodes = (int *)malloc(sizeof(int) * 2 * (G ? G : 1));
        int *edges = NULL, epn[3];
        uncapped_pair(bits0, bits0 + (size_t)(S - 1) * c, G, nodes, &edges, epn);
        snprintf(path, sizeof(path), "%s_EdgeList.txt", pathwayName);
        write_edgelist(path, genes, G, nodes, edges, epn, prior, class1, class2);
        free(nodes);
        free(edges);
      }
    }
    printf("\nPathway finished.\n\n");
    free(js); free(cls); free(t1); free(t2); free(bits0); free(f_ids);
    free(f_nodes); free(f_edges); free(f_epn); free(f_lval);
    free(prior); free(setIdx); free(genes); free(gsGenes); free(line);
  }
  fclose(fpGeneSet);
  fclose(fpResults);
  return 0;
}
