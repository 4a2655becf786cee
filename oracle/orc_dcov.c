/*
 * dcov.gamma, the gamma tail and the HSIC/permutation statistics for the CPU
 * oracle (test infrastructure only; never linked into the product).
 *
 * Follows:
 *   kpcalg/R/dcovgamma.R:56-82 (kpcalg dcov.gamma semantics),
 *   fastkpc/src/legacy_dcov_gamma_cpp.cpp:60-78 (top-|lambda| selection with
 *     higher-index tie break), :135 (centring after selection),
 *     :283-300 (weighted cross sum), :420-455 (moments, variance operator
 *     order, p = 1 - pgamma(nV2, shape, scale)).
 * RSpectra eigs (largest magnitude, tol 1e-10) is replaced by an exact
 * eigendecomposition (Householder tridiagonalisation, implicit QL values,
 * inverse-iteration vectors): kpcalg's truncated statistic is defined by the
 * exact top-k eigenpairs, RSpectra approximates them to ~1e-10.
 */
#include "kpc_oracle.h"
#include "orc_internal.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

double orc_r_mean(const double* x, int n, int stride) {
  long double s = 0.0L;
  for (int i = 0; i < n; ++i) s += x[(size_t)i * stride];
  s /= (long double)n;
  long double c = 0.0L;
  for (int i = 0; i < n; ++i) c += (long double)x[(size_t)i * stride] - s;
  s += c / (long double)n;
  return (double)s;
}

/* ------------------------------------------------------------ pgamma */
/* Lower regularised incomplete gamma P(a, x): power series when x < a + 1,
 * otherwise 1 - Q with Q from the Legendre continued fraction evaluated by the
 * modified Lentz method (Numerical Recipes 6.2; Press et al. 2007). */
double orc_pgamma_lower(double x, double a) {
  if (!(a > 0.0) || isnan(x)) return NAN;
  if (x <= 0.0) return 0.0;
  if (isinf(x)) return 1.0;
  double lpre = -x + a * log(x) - lgamma(a);
  if (x < a + 1.0) {
    double ap = a, del = 1.0 / a, sum = del;
    for (int i = 0; i < 100000; ++i) {
      ap += 1.0;
      del *= x / ap;
      sum += del;
      if (fabs(del) < fabs(sum) * 1e-17) break;
    }
    return sum * exp(lpre);
  }
  const double tiny = 1e-300;
  double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int i = 1; i < 100000; ++i) {
    double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (fabs(d) < tiny) d = tiny;
    c = b + an / c;
    if (fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < 1e-17) break;
  }
  double q = exp(lpre) * h;
  return 1.0 - q;
}

/* kpcalg p-value: 1 - pgamma(q, shape=alpha, rate=1/beta). */
double orc_gamma_pvalue(double stat, double mean, double var) {
  double alpha = (mean * mean) / var;
  double beta = var / mean;
  double rate = 1.0 / beta;
  double scale = 1.0 / rate;
  return 1.0 - orc_pgamma_lower(stat / scale, alpha);
}

/* ------------------------------------------------------ dcov component */
/* Top-k |lambda| eigenpairs of D_ij = |x_i - x_j| (index ignored, as kpcalg's
 * dist()), vectors centred by column means after selection. */
int orc_dcov_component(const double* x, int n, int k, double* values,
                       double* centered, double* dsum, double* moment) {
  if (n < 6 || k <= 0 || k >= n) return -1;
  double* a = (double*)malloc(sizeof(double) * (size_t)n * n);
  double* d = (double*)malloc(sizeof(double) * (size_t)n);
  double* e = (double*)malloc(sizeof(double) * (size_t)n);
  double* tau = (double*)malloc(sizeof(double) * (size_t)n);
  double* w = (double*)malloc(sizeof(double) * (size_t)n);
  double* dd = (double*)malloc(sizeof(double) * (size_t)n);
  double* ee = (double*)malloc(sizeof(double) * (size_t)n);
  int* order = (int*)malloc(sizeof(int) * (size_t)n);
  long double total = 0.0L;
  for (int j = 0; j < n; ++j) {
    a[(size_t)j * n + j] = 0.0;
    for (int i = j + 1; i < n; ++i) {
      double v = fabs(x[i] - x[j]);
      a[(size_t)j * n + i] = v;
      a[(size_t)i * n + j] = v;
      total += 2.0L * v;
    }
  }
  *dsum = (double)total;
  orc_tridiag_reduce(a, n, n, d, e, tau);
  memcpy(dd, d, sizeof(double) * n);
  memcpy(ee, e, sizeof(double) * (n - 1));
  memcpy(w, d, sizeof(double) * n);
  if (orc_tridiag_ql(w, ee, n, NULL, n) != 0) return -2;
  /* sort ascending so the index tie-break refers to ascending order */
  for (int i = 0; i < n; ++i) order[i] = i;
  for (int i = 1; i < n; ++i) {
    double v = w[i];
    int j = i - 1;
    while (j >= 0 && w[j] > v) { w[j + 1] = w[j]; --j; }
    w[j + 1] = v;
  }
  /* select top-k by |lambda|, tie -> higher index (ascending order index) */
  for (int i = 0; i < n; ++i) order[i] = i;
  for (int i = 1; i < n; ++i) {
    int oi = order[i];
    int j = i - 1;
    while (j >= 0) {
      int oj = order[j];
      double la = fabs(w[oi]), lb = fabs(w[oj]);
      int before = la > lb || (la == lb && oi > oj);
      if (!before) break;
      order[j + 1] = order[j];
      --j;
    }
    order[j + 1] = oi;
  }
  unsigned seed = 12345u;
  double tnorm = 0.0;
  for (int i = 0; i < n; ++i) {
    double s = fabs(d[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
    if (s > tnorm) tnorm = s;
  }
  double* prev = (double*)malloc(sizeof(double) * (size_t)n * k);
  for (int c = 0; c < k; ++c) {
    double lam = w[order[c]];
    values[c] = lam;
    int np = 0;
    for (int q = 0; q < c; ++q) {
      if (fabs(values[q] - lam) <= 1e-3 * tnorm) {
        memcpy(prev + (size_t)np * n, centered + (size_t)q * n, sizeof(double) * n);
        ++np;
      }
    }
    orc_tridiag_invit(d, e, n, lam, prev, np, centered + (size_t)c * n, &seed);
  }
  orc_tridiag_backtransform(a, n, n, tau, centered, k);
  for (int c = 0; c < k; ++c) {
    double* v = centered + (size_t)c * n;
    double m = 0.0;
    for (int i = 0; i < n; ++i) m += v[i];
    m /= n;
    for (int i = 0; i < n; ++i) v[i] -= m;
  }
  *moment = orc_weighted_cross_sum(centered, values, centered, values, n, k);
  free(a); free(d); free(e); free(tau); free(w); free(dd); free(ee); free(order); free(prev);
  return 0;
}

double orc_weighted_cross_sum(const double* cx, const double* lx, const double* cy,
                              const double* ly, int n, int k) {
  double total = 0.0;
  for (int j = 0; j < k; ++j) {
    for (int i = 0; i < k; ++i) {
      double s = 0.0;
      for (int r = 0; r < n; ++r) s += cx[(size_t)i * n + r] * cy[(size_t)j * n + r];
      total += lx[i] * ly[j] * s * s;
    }
  }
  return total;
}

void orc_dcov_pair_from_components(int n, int k, const double* cx, const double* lx,
                                   double dsum_x, double mom_x, const double* cy,
                                   const double* ly, double dsum_y, double mom_y,
                                   double out[5]) {
  double nd = (double)n;
  double nV2 = orc_weighted_cross_sum(cx, lx, cy, ly, n, k) / nd;
  double mean = (dsum_x / (nd * nd)) * (dsum_y / (nd * nd));
  double vf = 2.0 * (nd - 4.0) * (nd - 5.0) / nd / (nd - 1.0) / (nd - 2.0) / (nd - 3.0);
  double var = vf * mom_x * mom_y / pow(nd, 4.0) * pow(nd, 2.0);
  out[0] = orc_gamma_pvalue(nV2, mean, var);
  out[1] = nV2;
  out[2] = mean;
  out[3] = var;
  out[4] = nV2;
}

int orc_dcov_gamma(const double* x, const double* y, int n, int num_col, double out[5]) {
  double* cx = (double*)malloc(sizeof(double) * (size_t)n * num_col);
  double* cy = (double*)malloc(sizeof(double) * (size_t)n * num_col);
  double* lx = (double*)malloc(sizeof(double) * num_col);
  double* ly = (double*)malloc(sizeof(double) * num_col);
  double sx, sy, mx, my;
  int rc = orc_dcov_component(x, n, num_col, lx, cx, &sx, &mx);
  if (rc == 0) rc = orc_dcov_component(y, n, num_col, ly, cy, &sy, &my);
  if (rc == 0) orc_dcov_pair_from_components(n, num_col, cx, lx, sx, mx, cy, ly, sy, my, out);
  free(cx); free(cy); free(lx); free(ly);
  return rc;
}
