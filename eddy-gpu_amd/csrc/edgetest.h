// edgetest.h - the chi-square edge decision of one 3x3 tally (kernel1.cu:10-203,
// 420-431), shared by the gfx950 kernels and a host build used by the tests.
//
// eddy_edge_decision() returns exactly what the reference's decision computes
// (chi-square with its Yates rule, AS 147 series, Bonferroni, prior
// threshold), but skips the AS 147 series when the outcome is already certain
// from x = chi2/2 alone.  For dof 1, 2, 4 (p = 1/2, 1, 2) the upper tail
// Q(p,x) = 1 - P(p,x) has closed forms (erfc(sqrt x), e^-x, e^-x (1+x)), so the
// host computes per gene set (its Bonferroni factor c) the x intervals where:
//   x <= xlo       Q(x) >= thr (1 + 1e-5): the series (which only
//                  underestimates P, by truncation at 1e-9 of the sum, and
//                  overestimates it by at most ~6e-14 relative through
//                  rounding) cannot bring 1 - P below thr: no edge;
//   xhi <= x < xclo  Q(x) <= (thr - 5e-9)(1 - 1e-5): the truncation error
//                  (measured <= 1.4e-9, bounded by 2e-9) cannot lift 1 - P
//                  to thr, and x is below AS 147's underflow cut: an edge;
//   x > xchi       AS 147's underflow cut p ln x - lgamma(p+1) - x < ln 1e-37
//                  holds with margin: P = 0, p-value 1, no edge;
// and runs the exact computation everywhere else (within the margins of a
// threshold, and for every other dof).  thr = max(alphaEdge, alphaEdgePrior
// if the pair has a prior edge) / c with Bonferroni, without it thr = alpha.
//
// The fast forms (eddy_edge_decision_fast, eddy_loo_mask_fast) compute x from
// the margins' reciprocals instead of the reference's 18 divisions per tally:
// the same x within 1e-10 (1 + x) absolute (error analysis at
// eddy_half_chi2_fast; measured in the tests), so they apply the intervals
// above shrunk by a guard of 1e-9 (1 + x) and run the exact decision on the
// tally itself wherever x falls inside a guard band.
#pragma once
#include "fpmath.h"

typedef struct {
  double aE, aEP;           // alphaEdge, alphaEdgePrior
  double lg15, lg2, lg3;    // lgamma(1.5), lgamma(2), lgamma(3) (AS 147's lgamma(p+1))
  int p_adjust;             // Bonferroni on/off
  int shortcut;             // 0: always the exact series
  double xclo[3], xchi[3];  // underflow-cut interval per p = 1/2, 1, 2
} eddy_edge_params;

typedef struct {
  double xlo[2][3];         // [prior][p index]; < 0 disables
  double xhi[2][3];         // > xclo disables
} eddy_edge_bounds;

// AS 147: kernel1.cu:10-74
EDDY_HD double eddy_gammds(double x, double p, double lgp1) {
  if (x != x) return 0.0;  // unreachable for integer tallies; keeps the series finite
  if (x <= 0.0) return 0.0;
  if (p <= 0.0) return 0.0;
  double arg = p * eddy_log(x) - lgp1 - x;
  if (arg < eddy_log(1.0e-37)) return 0.0;
  double f = eddy_exp(arg);
  if (f == 0.0) return 0.0;
  double c = 1.0, value = 1.0, a = p;
  for (;;) {
    a = a + 1.0;
    c = c * x / a;
    value = value + c;
    if (c <= 1.0e-9 * value) break;
  }
  return value * f;
}

// chi-square and dof: kernel1.cu:121-203 (row-major sum, Yates on all cells
// when any expected count of a non-empty row/column is below 5)
EDDY_HD double eddy_chi2(const int *t, int *dof) {
  int ex[6] = {0, 0, 0, 0, 0, 0};
  for (int r = 0; r < 3; r++)
    for (int k = 0; k < 3; k++) {
      ex[k] += t[r * 3 + k];      // column sums
      ex[3 + r] += t[r * 3 + k];  // row sums
    }
  double divisor = (double)(ex[0] + ex[1] + ex[2]);
  double expected[9];
  for (int r = 0; r < 3; r++)
    for (int k = 0; k < 3; k++) expected[r * 3 + k] = (double)ex[k] * (double)ex[r + 3] / divisor;
  double yates = 0.0;
  for (int r = 0; r < 3; r++)
    for (int k = 0; k < 3; k++)
      if (expected[r * 3 + k] < 5 && ex[k] && ex[r + 3]) yates = 0.5;
  double chi = 0.0;
  int dofm = 0, dofn = 0;
  for (int ii = 0; ii < 3; ii++) {
    dofm += ex[ii] == 0;
    dofn += ex[ii + 3] == 0;
    for (int jj = 0; jj < 3; jj++) {
      if (ex[jj] * ex[ii + 3] != 0) {
        double d = eddy_fabs((double)t[ii * 3 + jj] - expected[ii * 3 + jj]) - yates;
        chi = chi + d * d / expected[ii * 3 + jj];
      }
    }
  }
  *dof = ((3 - dofm) - 1) * ((3 - dofn) - 1);
  return chi;
}

// exact decision: kernel1.cu:420-431
EDDY_HD int eddy_edge_exact(double chi, int dof, int prior, const eddy_edge_params *ep, double cpairs) {
  double lgp1 = dof == 1 ? ep->lg15 : dof == 2 ? ep->lg2 : ep->lg3;
  double pv = 1.0 - eddy_gammds(chi / 2.0, ((double)dof) / 2.0, lgp1);
  if (ep->p_adjust) {
    double adj = pv * cpairs;
    pv = adj < 1.0 ? adj : 1.0;
  }
  return (pv < ep->aE || (prior == 1 && pv < ep->aEP)) ? 1 : 0;
}

// v[0], v[1] or v[2] by constant indices and selects (no dynamic indexing,
// so a caller's local copy of the bounds stays in registers)
EDDY_HD double eddy_pick3(const double *v, bool p1, bool p2) { return p1 ? v[0] : p2 ? v[1] : v[2]; }

EDDY_HD int eddy_edge_decision(const int *t, int prior, const eddy_edge_params *ep, const eddy_edge_bounds *eb,
                               double cpairs) {
  int dof;
  double chi = eddy_chi2(t, &dof);
  if (ep->shortcut && (dof == 1 || dof == 2 || dof == 4)) {
    const bool p1 = dof == 1, p2 = dof == 2, pr = prior == 1;
    const double x = chi / 2.0;
    const double xlo = pr ? eddy_pick3(eb->xlo[1], p1, p2) : eddy_pick3(eb->xlo[0], p1, p2);
    const double xhi = pr ? eddy_pick3(eb->xhi[1], p1, p2) : eddy_pick3(eb->xhi[0], p1, p2);
    if (x <= xlo) return 0;
    if (x > eddy_pick3(ep->xchi, p1, p2)) return 0;
    if (x >= xhi && x < eddy_pick3(ep->xclo, p1, p2)) return 1;
  }
  return eddy_edge_exact(chi, dof, prior, ep, cpairs);
}

// ---------------------------------------------------------------------------
// fast forms
// ---------------------------------------------------------------------------
// The reciprocals of three margins from one division: with P the product of
// the nonzero ones (exact: three integers <= 20000 multiply to < 2^53) and
// inv = 1/P rounded once, r_i = (P / v_i) inv, where P / v_i is an exact
// integer product, is within 1 ulp of 1 / v_i; zero margins get 0.  With
// `minus1` the reciprocals of v_i - 1 (0 where v_i < 2).
EDDY_HD void eddy_recip3(const int *v, double *r, int minus1) {
  double w[3], P = 1.0;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const int x = minus1 ? v[i] - 1 : v[i];
    w[i] = x > 0 ? (double)x : 0.0;
    if (x > 0) P = P * w[i];
  }
  const double inv = 1.0 / P;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double o = 1.0;
#pragma unroll
    for (int j = 0; j < 3; j++)
      if (j != i && w[j] > 0.0) o = o * w[j];
    r[i] = w[i] > 0.0 ? o * inv : 0.0;
  }
}

EDDY_HD double eddy_sq(int v) { return (double)(v * v); }

// x = chi2 / 2 and dof of tally t with row sums R, column sums K, total n,
// from iR[i] = 1/R[i], iK[j] = 1/K[j] (0 for empty sums, within 1.5 ulp).  Yates' rule is decided exactly in integers: (double)K*R/n < 5
// holds iff R*K < 5n, since a quotient of integers <= 2^22 that is not 5
// differs from 5 by at least 1/n, far more than one rounding.
//  - Without Yates' correction, the identity sum (O - E)^2 / E =
//    n (sum O^2 / (R K) - 1) (empty rows and columns left out): each product
//    O^2 iK iR is within 5 ulp, their sum S <= 3 within 12 ulp of itself, so
//    |x - x_exact| <= 0.5 n 36 u + 2 u x < 4e-11 + 2.3e-16 x (n <= 20000).
//  - With it, per cell (|O - E| - y)^2 n iR iK, which differs from the
//    reference's d*d/E by a few ulp of itself plus 2|d| * 3 ulp(E) <=
//    2 n 6.7e-16, so |x - x_ref| <= 9 n 6.7e-16 + 1e-15 x < 1.3e-10 + 1e-15 x
//    (n <= 20000).
EDDY_HD double eddy_half_chi2_fast(const int *t, const int *R, const int *K, int n, const double *iR,
                                   const double *iK, int *dof) {
  int zr = 0, zk = 0, yates = 0;
  for (int i = 0; i < 3; i++) {
    zr += R[i] == 0;
    zk += K[i] == 0;
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      if (R[i] && K[j] && R[i] * K[j] < 5 * n) yates = 1;
  *dof = (2 - zk) * (2 - zr);
  if (!yates) {
    double S = 0.0;
    for (int i = 0; i < 3; i++) {
      double r = 0.0;
      for (int j = 0; j < 3; j++) r = r + eddy_sq(t[i * 3 + j]) * iK[j];
      S = S + iR[i] * r;
    }
    return 0.5 * (double)n * (S - 1.0);
  }
  const double y = 0.5, dn = (double)n, in = 1.0 / dn;
  double chi = 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      if (R[i] && K[j]) {
        const double rk = (double)(R[i] * K[j]);
        const double d = eddy_fabs((double)t[i * 3 + j] - rk * in) - y;
        chi += d * d * dn * (iR[i] * iK[j]);
      }
  return 0.5 * chi;
}

// certain outcome of x (1 / 0) from the intervals shrunk by the guard, or -1
EDDY_HD int eddy_edge_certain(double x, int dof, int prior, const eddy_edge_params *ep, const eddy_edge_bounds *eb) {
  if (!ep->shortcut || !(dof == 1 || dof == 2 || dof == 4)) return -1;
  // constant indices and selects only: the bounds stay in registers when
  // the caller holds them in a local copy
  const bool p1 = dof == 1, p2 = dof == 2, pr = prior == 1;
  const double xlo = pr ? eddy_pick3(eb->xlo[1], p1, p2) : eddy_pick3(eb->xlo[0], p1, p2);
  const double xhi = pr ? eddy_pick3(eb->xhi[1], p1, p2) : eddy_pick3(eb->xhi[0], p1, p2);
  const double xchi = eddy_pick3(ep->xchi, p1, p2);
  const double xclo = eddy_pick3(ep->xclo, p1, p2);
  const double gd = 1e-9 * (1.0 + x);
  if (x <= xlo - gd) return 0;
  if (x > xchi + gd) return 0;
  if (x >= xhi + gd && x < xclo - gd) return 1;
  return -1;
}



// ---------------------------------------------------------------------------
// single-precision screens
// ---------------------------------------------------------------------------
// Most tallies lie far from every threshold, so their decision is already
// certain from x computed in single precision, with no FP64 division.  The
// screens below decide only where an error bound of that x keeps them inside
// a certain region; everywhere else they return -1 and the caller takes the
// double-precision path above, so decisions never change.  (The
// leave-one-out kernel keeps the double-precision forms: a screen there
// changes its register allocation more than it saves.)

// 1 / v in single precision: the hardware reciprocal (within 1 ulp) on the
// GPU, the correctly rounded quotient on the host.
EDDY_HD float eddy_rcp32(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(v);
#else
  return 1.0f / v;
#endif
}

// S = sum O^2 / (R K) over the nonempty rows and columns, in single
// precision.  u = 2^-24: O^2 <= 4e8 rounds once (u), 1/K and 1/R are within
// 1 ulp (2u), each product and sum rounds once, so every term of a row sum is
// within 4u, the row sum of three nonnegative terms within 6u, and with 1/R
// (2u) and the product and sum of each of the three steps (2u each) S is
// within 14u of itself; S <= 3 (S <= the number of nonempty rows), so
// |S - S_exact| <= 42u < 2.6e-6.
EDDY_HD float eddy_S32(const int *t, const int *R, const int *K) {
  float iK[3], S = 0.0f;
  for (int j = 0; j < 3; j++) iK[j] = K[j] > 0 ? eddy_rcp32((float)K[j]) : 0.0f;
  for (int i = 0; i < 3; i++) {
    const float iR = R[i] > 0 ? eddy_rcp32((float)R[i]) : 0.0f;
    float r = 0.0f;
    for (int j = 0; j < 3; j++) r = r + (float)(t[i * 3 + j] * t[i * 3 + j]) * iK[j];
    S = S + iR * r;
  }
  return S;
}
// An upper bound of |x32 - x_exact| for x32 = fl(0.5 n fl(S32 - 1)): the
// subtraction adds u |S - 1| <= 2u, the product with n one rounding of x, so
// |x32 - x| <= 0.5 n 44u + u |x| < 1.4e-6 n + 6e-8 |x|; the guard doubles it.
EDDY_HD double eddy_x32_guard(int n, double x) { return 3e-6 * (double)(n + 1) + 1e-6 * eddy_fabs(x); }

// The screen's thresholds on x32 for tables of n samples: per prior flag and
// dof (1, 2, 4) {A, B, C, D}; x32 <= A or x32 > B: certainly no edge,
// C <= x32 < D: certainly an edge.  They fold the guard g = 3e-6 (n + 1) +
// 1e-6 x (x >= 0) into the interval ends (x + g <= xlo  <=>  x <= (xlo - c) /
// (1 + 1e-6) with c = 3e-6 (n + 1), and alike) and are rounded toward deciding
// less: each is computed in double, moved 1e-12 of itself the safe way (more
// than the double rounding of the quotient) and rounded to float in the same
// direction, so x32 against them implies the guarded comparisons in real
// arithmetic.  n < 0 (or the shortcut off) decides nothing.
struct eddy_screen_thr {
  float v[2][3][4];
  int n;
};
EDDY_HD float eddy_f32_down(double v) {
  v = v > 0.0 ? v * (1.0 - 1e-12) : v * (1.0 + 1e-12);
  float f = (float)v;
  if ((double)f > v) f = nextafterf(f, -INFINITY);
  return f;
}
EDDY_HD float eddy_f32_up(double v) {
  v = v > 0.0 ? v * (1.0 + 1e-12) : v * (1.0 - 1e-12);
  float f = (float)v;
  if ((double)f < v) f = nextafterf(f, INFINITY);
  return f;
}
EDDY_HD void eddy_screen_thresholds(const eddy_edge_params *ep, const eddy_edge_bounds *eb, int n,
                                    eddy_screen_thr *th) {
  const double c = 3e-6 * (double)(n + 1), up = 1.0 + 1e-6, dn = 1.0 - 1e-6;
  for (int pr = 0; pr < 2; pr++)
    for (int d = 0; d < 3; d++) {
      float *T = th->v[pr][d];
      T[0] = eddy_f32_down((eb->xlo[pr][d] - c) / up);
      T[1] = eddy_f32_up((ep->xchi[d] + c) / dn);
      T[2] = eddy_f32_up((eb->xhi[pr][d] + c) / dn);
      T[3] = eddy_f32_down((ep->xclo[d] - c) / up);
    }
  th->n = ep->shortcut && n > 0 ? n : -1;
}

// The full test's decision from x in single precision: 1 / 0 when certain,
// -1 otherwise (Yates' rule, other dofs, x near an interval end, or tables of
// another sample count than the thresholds').  x32 is clamped at 0: the exact
// x is not negative, so the clamp only moves x32 toward it.
EDDY_HD int eddy_edge_screen32t(const int *t, const int *R, const int *K, int n, int prior,
                                const eddy_screen_thr *th) {
  if (n != th->n) return -1;
  int zr = 0, zk = 0, rmin = 0x7fffffff, kmin = 0x7fffffff;
  for (int i = 0; i < 3; i++) {
    zr += R[i] == 0;
    zk += K[i] == 0;
    if (R[i] > 0 && R[i] < rmin) rmin = R[i];
    if (K[i] > 0 && K[i] < kmin) kmin = K[i];
  }
  // Yates' rule: some cell of a nonempty row and column expects fewer than 5
  // (the smallest such product is the product of the smallest margins)
  const int dof = (2 - zk) * (2 - zr);
  if (rmin * kmin < 5 * n || !(dof == 1 || dof == 2 || dof == 4)) return -1;
  float x32 = 0.5f * ((float)n * (eddy_S32(t, R, K) - 1.0f));
  x32 = x32 > 0.0f ? x32 : 0.0f;
  // every lane reads the same rows (broadcast loads when th is in LDS), then
  // picks its own prior flag and dof
  const bool p1 = dof == 1, p2 = dof == 2, pr = prior == 1;
  float T[4];
  for (int k = 0; k < 4; k++) {
    const float a = p1 ? th->v[0][0][k] : p2 ? th->v[0][1][k] : th->v[0][2][k];
    const float b = p1 ? th->v[1][0][k] : p2 ? th->v[1][1][k] : th->v[1][2][k];
    T[k] = pr ? b : a;
  }
  if (x32 <= T[0] || x32 > T[1]) return 0;
  if (x32 >= T[2] && x32 < T[3]) return 1;
  return -1;
}

// The same with the thresholds made here (one table's screen)
EDDY_HD int eddy_edge_screen32(const int *t, const int *R, const int *K, int n, int prior,
                               const eddy_edge_params *ep, const eddy_edge_bounds *eb) {
  eddy_screen_thr th;
  eddy_screen_thresholds(ep, eb, n, &th);
  return eddy_edge_screen32t(t, R, K, n, prior, &th);
}

// eddy_edge_decision() with the fast x
EDDY_HD int eddy_edge_decision_fast(const int *t, int prior, const eddy_edge_params *ep, const eddy_edge_bounds *eb,
                                    double cpairs) {
  int R[3], K[3];
  for (int i = 0; i < 3; i++) {
    R[i] = t[i * 3] + t[i * 3 + 1] + t[i * 3 + 2];
    K[i] = t[i] + t[3 + i] + t[6 + i];
  }
  const int n = R[0] + R[1] + R[2];
  if (n > 0) {
    double iR[3], iK[3];
    eddy_recip3(R, iR, 0);
    eddy_recip3(K, iK, 0);
    int dof;
    const double x = eddy_half_chi2_fast(t, R, K, n, iR, iK, &dof);
    const int c = eddy_edge_certain(x, dof, prior, ep, eb);
    if (c >= 0) return c;
  }
  return eddy_edge_decision(t, prior, ep, eb, cpairs);
}

// The screen alone: 1 / 0 when certain, -1 when the double-precision path
// must decide
EDDY_HD int eddy_edge_screened_only(const int *t, int prior, const eddy_screen_thr *th) {
  int R[3], K[3];
  for (int i = 0; i < 3; i++) {
    R[i] = t[i * 3] + t[i * 3 + 1] + t[i * 3 + 2];
    K[i] = t[i] + t[3 + i] + t[6 + i];
  }
  return eddy_edge_screen32t(t, R, K, R[0] + R[1] + R[2], prior, th);
}

// eddy_edge_decision_fast() behind the single-precision screen: the count
// kernel's full tests, where the two FP64 divisions per test are most of the
// test's cost and the screen decides nearly every table; th holds the
// thresholds for the class's sample count
EDDY_HD int eddy_edge_decision_screened(const int *t, int prior, const eddy_edge_params *ep,
                                        const eddy_edge_bounds *eb, double cpairs, const eddy_screen_thr *th) {
  int R[3], K[3];
  for (int i = 0; i < 3; i++) {
    R[i] = t[i * 3] + t[i * 3 + 1] + t[i * 3 + 2];
    K[i] = t[i] + t[3 + i] + t[6 + i];
  }
  const int c32 = eddy_edge_screen32t(t, R, K, R[0] + R[1] + R[2], prior, th);
  if (c32 >= 0) return c32;
  return eddy_edge_decision_fast(t, prior, ep, eb, cpairs);
}

// x = chi2 / 2 of the tally t minus one count in cell (a, b), t[a*3+b] > 0,
// without Yates' correction, from the identity
//   sum (O - E)^2 / E = n (sum O^2 / (R K) - 1)     (E = R K / n, empty rows
//                                                    and columns left out):
//   x = (n - 1) / 2 * (sum_i 1/R'_i sum_j O'_ij^2 / K'_j - 1).
// T[i] = sum_{j != b} t_ij^2 / K_j and U[i] = T[i] + t_ib^2 / (K_b - 1) are
// shared by the three cells of column b; row a takes 1/(R_a - 1) and
// (t_ab - 1)^2.  Error: every product is within 5 ulp, the sum S of at most
// nine nonnegative terms (S <= 3) within 12 ulp of itself, so
// |x - x_exact| <= 0.5 (n - 1) 36 u + 2 u x < 4.1e-12 + 2.3e-16 x for
// n <= 2046 (4e-11 + 2.3e-16 x for n <= 20000; u = 2^-53): with the reference
// formula's own rounding this stays far inside the 1e-9 (1 + x) guard of
// eddy_edge_certain.
EDDY_HD double eddy_loo_half_chi2(const int *t, int n, int a, int b, const double *T, const double *U,
                                  const double *iR, const double *iRm, const double *iKm) {
  const double W = T[a] + eddy_sq(t[a * 3 + b] - 1) * iKm[b];
  double S = 0.0;
#pragma unroll
  for (int i = 0; i < 3; i++) S = S + (i == a ? iRm[i] * W : iR[i] * U[i]);
  return 0.5 * (double)(n - 1) * (S - 1.0);
}

// x and dof of the tally t minus one count in cell (a, b) (t[a*3+b] > 0,
// n > 1), given t's margins R, K, their reciprocal pairs and column b's T, U:
// the identity above, or, when an expected count of the reduced tally is
// below 5 (Yates' rule, decided exactly in integers), the per-cell form.
EDDY_HD double eddy_loo_x(const int *t, const int *R, const int *K, int n, int a, int b, const double *T,
                          const double *U, const double *iR, const double *iRm, const double *iK,
                          const double *iKm, int *dof) {
  int R1[3], K1[3], zr = 0, zk = 0, mr = 1 << 30, mk = 1 << 30;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    R1[i] = R[i] - (i == a ? 1 : 0);
    K1[i] = K[i] - (i == b ? 1 : 0);
    zr += R1[i] == 0;
    zk += K1[i] == 0;
    if (R1[i] > 0 && R1[i] < mr) mr = R1[i];
    if (K1[i] > 0 && K1[i] < mk) mk = K1[i];
  }
  *dof = (2 - zk) * (2 - zr);
  if (mr * mk >= 5 * (n - 1)) return eddy_loo_half_chi2(t, n, a, b, T, U, iR, iRm, iKm);
  int u[9], d;
  double iR1[3], iK1[3];
#pragma unroll
  for (int i = 0; i < 9; i++) u[i] = t[i] - (i == a * 3 + b ? 1 : 0);
#pragma unroll
  for (int i = 0; i < 3; i++) {
    iR1[i] = i == a ? iRm[i] : iR[i];
    iK1[i] = i == b ? iKm[i] : iK[i];
  }
  return eddy_half_chi2_fast(u, R1, K1, n - 1, iR1, iK1, &d);
}

// margins of t, their reciprocal pairs, n
EDDY_HD int eddy_loo_margins(const int *t, int *R, int *K, double *iR, double *iRm, double *iK, double *iKm) {
  for (int i = 0; i < 3; i++) {
    R[i] = t[i * 3] + t[i * 3 + 1] + t[i * 3 + 2];
    K[i] = t[i] + t[3 + i] + t[6 + i];
  }
  eddy_recip3(R, iR, 0);
  eddy_recip3(R, iRm, 1);
  eddy_recip3(K, iK, 0);
  eddy_recip3(K, iKm, 1);
  return R[0] + R[1] + R[2];
}

// T[i] = sum_{j != b} t_ij^2 / K_j and U[i] = T[i] + t_ib^2 / (K_b - 1)
EDDY_HD void eddy_loo_column(const int *t, int b, const double *iK, const double *iKm, double *T, double *U) {
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++)
      if (j != b) s = s + eddy_sq(t[i * 3 + j]) * iK[j];
    T[i] = s;
    U[i] = s + eddy_sq(t[i * 3 + b]) * iKm[b];
  }
}

// The margin conditions of the all-at-once argument (below): every nonzero margin >= 2,
// (Rmin - 1)(Kmin - 1) >= 5 (n - 1) and dof 1, 2 or 4; then *e = an upper
// bound of 1/(Rmin - 1) + 1/(Kmin - 1) (one division of exact integers,
// correctly rounded; the 1 + 1e-12 factor below covers its rounding) and *dof
// are set.  1 if they hold.
EDDY_HD int eddy_loo_all_margins(const int *R, const int *K, int n, double *e, int *dof) {
  int rmin = 1 << 30, kmin = 1 << 30, zr = 0, zk = 0;
  for (int i = 0; i < 3; i++) {
    if (R[i] == 0) zr++;
    else if (R[i] < rmin) rmin = R[i];
    if (K[i] == 0) zk++;
    else if (K[i] < kmin) kmin = K[i];
  }
  if (n < 2 || rmin < 2 || kmin < 2) return 0;
  const long long pm = (long long)(rmin - 1) * (long long)(kmin - 1);
  if (pm < 5LL * (n - 1)) return 0;
  *dof = (2 - zk) * (2 - zr);
  if (!(*dof == 1 || *dof == 2 || *dof == 4)) return 0;
  *e = (double)(rmin + kmin - 2) / (double)pm;
  return 1;
}

// Every leave-one-out decision of a tally at once, when its statistic is far
// enough from the thresholds: 1 all nine are "edge", 0 all are "no edge", -1
// unknown.  With every nonzero margin >= 2, removing one sample from cell
// (a, b) keeps the zero rows and columns (so the dof) and changes
// S = sum O^2 / (R K) by at most e = 1/(R_a - 1) + 1/(K_b - 1): the other
// cells of row a and of column b gain at most 1/(R_a - 1) and 1/(K_b - 1)
// (O_aj^2 / K_j <= O_aj), and the cell itself loses at most e
// ((O-1)/(R-1) = O/R - (1 - O/R)/(R - 1)).  With x = n (S - 1) / 2, the
// leave-one-out x' = (n - 1)(S' - 1) / 2 then lies in
// [x - (n - 1) e / 2 - x / n, x + (n - 1) e / 2], e taken at the smallest
// margins.  (Rmin - 1)(Kmin - 1) >= 5 (n - 1) keeps Yates' rule out of every
// leave-one-out tally.  The fast x is within 1e-9 (1 + x) of the exact value,
// and each interval end gets the same guard, so an answer here is what the
// exact decision of each leave-one-out tally gives.
EDDY_HD int eddy_loo_all_certain(const int *R, const int *K, int n, const double *iR, const double *iK,
                                 const int *t, int prior, const eddy_edge_params *ep, const eddy_edge_bounds *eb) {
  double e;
  int dof;
  if (!ep->shortcut || !eddy_loo_all_margins(R, K, n, &e, &dof)) return -1;
  double S = 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) S = S + eddy_sq(t[i * 3 + j]) * iR[i] * iK[j];
  const double x = 0.5 * (double)n * (S - 1.0);
  const double d = 0.5 * (double)(n - 1) * e * (1.0 + 1e-12);
  const double xg = x + 1e-9 * (1.0 + (x > 0.0 ? x : 0.0));  // >= the exact x
  const double up = xg + d;                                   // >= every leave-one-out x
  const double dn = x - 1e-9 * (1.0 + (x > 0.0 ? x : 0.0)) - d - xg / (double)n;  // <= every one
  const bool p1 = dof == 1, p2 = dof == 2, pr = prior == 1;
  const double xlo = pr ? eddy_pick3(eb->xlo[1], p1, p2) : eddy_pick3(eb->xlo[0], p1, p2);
  const double xhi = pr ? eddy_pick3(eb->xhi[1], p1, p2) : eddy_pick3(eb->xhi[0], p1, p2);
  const double xclo = eddy_pick3(ep->xclo, p1, p2);
  if (up <= xlo - 1e-9 * (1.0 + (xlo > 0.0 ? xlo : 0.0))) return 0;
  if (dn >= xhi + 1e-9 * (1.0 + xhi) && up <= xclo - 1e-9 * (1.0 + xclo)) return 1;
  return -1;
}

// The leave-one-out mask that a certain common decision `all` gives:
// nonzero cells take `all`, empty cells copy `full`.
EDDY_HD uint32_t eddy_loo_all_mask(const int *t, int all, int full) {
  uint32_t nz = 0;
  for (int c = 0; c < 9; c++) nz |= (t[c] > 0 ? 1u : 0u) << c;
  return (all ? nz : 0u) | (full ? (~nz & 0x1FFu) : 0u);
}

// The leave-one-out mask of t when the all-at-once argument decides it (two
// divisions: the margins' reciprocals): *m and 1; else 0.
EDDY_HD int eddy_loo_mask_try(const int *t, int full, int prior, const eddy_edge_params *ep,
                              const eddy_edge_bounds *eb, uint32_t *m) {
  int R[3], K[3];
  for (int i = 0; i < 3; i++) {
    R[i] = t[i * 3] + t[i * 3 + 1] + t[i * 3 + 2];
    K[i] = t[i] + t[3 + i] + t[6 + i];
  }
  double iR[3], iK[3];
  eddy_recip3(R, iR, 0);
  eddy_recip3(K, iK, 0);
  const int all = eddy_loo_all_certain(R, K, R[0] + R[1] + R[2], iR, iK, t, prior, ep, eb);
  if (all < 0) return 0;
  *m = eddy_loo_all_mask(t, all, full);
  return 1;
}

// Leave-one-out decisions of one tally, cell by cell: bit `cell` is the
// decision of t minus one count in `cell`; empty cells copy `full`.  Four
// divisions serve all nine leave-one-out tallies.  A reduced tally under
// Yates' rule (rare: an expected count below 5) or within a guard band of a
// threshold takes the exact decision after the unrolled pass.
EDDY_HD uint32_t eddy_loo_mask_cells(const int *t, int full, int prior, const eddy_edge_params *ep,
                                     const eddy_edge_bounds *eb, double cpairs, int *ntests) {
  int R[3], K[3];
  double iR[3], iK[3], iRm[3], iKm[3];
  const int n = eddy_loo_margins(t, R, K, iR, iRm, iK, iKm);
  uint32_t m = 0, unsure = 0;
  // unrolled: every array index below is a constant, so all tallies stay in
  // registers
#pragma unroll
  for (int b = 0; b < 3; b++) {
    double T[3], U[3];
    eddy_loo_column(t, b, iK, iKm, T, U);
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const int cell = a * 3 + b;
      if (t[cell] > 0) {
        (*ntests)++;
        int c = -1;
        if (n > 1) {
          int zr = 0, zk = 0, mr = 1 << 30, mk = 1 << 30;
#pragma unroll
          for (int i = 0; i < 3; i++) {
            const int r1 = R[i] - (i == a ? 1 : 0), k1 = K[i] - (i == b ? 1 : 0);
            zr += r1 == 0;
            zk += k1 == 0;
            if (r1 > 0 && r1 < mr) mr = r1;
            if (k1 > 0 && k1 < mk) mk = k1;
          }
          if (mr * mk >= 5 * (n - 1)) {  // no Yates correction: the identity
            const double x = eddy_loo_half_chi2(t, n, a, b, T, U, iR, iRm, iKm);
            c = eddy_edge_certain(x, (2 - zk) * (2 - zr), prior, ep, eb);
          }
        }
        if (c >= 0) m |= (uint32_t)c << cell;
        else unsure |= 1u << cell;
      } else {
        m |= (uint32_t)full << cell;
      }
    }
  }
  // tallies within a guard band of a threshold: the exact decision
  while (unsure) {
    const int cell = __builtin_ctz(unsure);
    unsure &= unsure - 1;
    int u[9];
#pragma unroll
    for (int i = 0; i < 9; i++) u[i] = t[i] - (i == cell ? 1 : 0);
    m |= (uint32_t)eddy_edge_decision(u, prior, ep, eb, cpairs) << cell;
  }
  return m;
}

// Leave-one-out decisions of one tally: all at once when the argument above
// decides them, else cell by cell.
EDDY_HD uint32_t eddy_loo_mask_fast(const int *t, int full, int prior, const eddy_edge_params *ep,
                                    const eddy_edge_bounds *eb, double cpairs, int *ntests) {
  uint32_t m;
  if (eddy_loo_mask_try(t, full, prior, ep, eb, &m)) return m;
  return eddy_loo_mask_cells(t, full, prior, ep, eb, cpairs, ntests);
}
