/*
 * HSIC tests for the CPU oracle (test infrastructure only).
 *
 * Incomplete Cholesky of the Gaussian kernel matrix (kernlab::inchol as used
 * by kpcalg/R/hsicgamma.R:60-66, restated after the reference kernel
 * fastkpc/src/cuda/full_cuda_ci_method_batch.cu:445-559): kernel
 * k(a, b) = exp(-sigma (a - b)^2) with sigma = 1 / sig; diagonal residues start
 * at 1, the first pivot is row 0, later pivots are the first maximum of the
 * residues; each step forward-solves the pivot's kernel column against the
 * triangle of earlier pivots, tau^2 = k(p, p) - |solve|^2, new column
 * (k(., p) - G solve) / tau, residues -= column^2; the loop runs while the
 * largest residue exceeds tol = 0.001 and rank < numCol, and stops early when
 * tau^2 is not positive.
 *
 * Component quantities (fastkpc/src/cuda/full_cuda_ci_method_batch.cu:545-658):
 * off-diagonal sum of K = G G^T (uncentred factor), column-centred factor
 * Gc = H G, self moment |Gc^T Gc|_F^2 (= trace((Ux Sx^2 Ux^T)^2) of kpcalg).
 *
 * hsic.gamma (kpcalg/R/hsicgamma.R:74-83): HSIC = |Gcx^T Gcy|_F^2 / n^2,
 * mean = (1 + mux muy - mux - muy) / n with mu = offdiag / (n (n - 1)),
 * var = 2 (n-4)(n-5) / (n (n-1)(n-2)(n-3)) * Mx * My / n^4,
 * p = 1 - pgamma(HSIC, shape = mean^2/var, rate = var/mean ^ -1).
 */
#include "kpc_oracle.h"
#include "orc_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

int orc_inchol(const double* x, int n, double sigma, double tol, int maxr, double* g,
               int* pivots, double* margin) {
  double* diag = (double*)malloc(sizeof(double) * n);
  double* col = (double*)malloc(sizeof(double) * n);
  double* upper = (double*)calloc((size_t)maxr * maxr, sizeof(double));
  double* solve = (double*)malloc(sizeof(double) * (maxr > 0 ? maxr : 1));
  int* piv = (int*)malloc(sizeof(int) * (maxr > 0 ? maxr : 1));
  for (int i = 0; i < n; ++i) diag[i] = 1.0;
  memset(g, 0, sizeof(double) * (size_t)n * maxr);
  double residue = 1.0, mg = 1e300;
  int pivot = 0, rank = 0;
  while (residue > tol && rank < maxr) {
    const double xp = x[pivot];
    for (int i = 0; i < n; ++i) {
      const double d = x[i] - xp;
      col[i] = exp(-sigma * d * d);
    }
    double tau2 = col[pivot];
    if (rank == 0) {
      upper[0] = sqrt(tau2);
    } else {
      for (int r = 0; r < rank; ++r) {
        double v = col[piv[r]];
        for (int q = 0; q < r; ++q) v -= upper[q + (size_t)maxr * r] * solve[q];
        solve[r] = v / upper[r + (size_t)maxr * r];
        tau2 -= solve[r] * solve[r];
        upper[r + (size_t)maxr * rank] = solve[r];
      }
    }
    if (!(tau2 > 0.0) || !isfinite(tau2)) break;
    const double tau = sqrt(tau2);
    upper[rank + (size_t)maxr * rank] = tau;
    for (int i = 0; i < n; ++i) {
      double proj = 0.0;
      for (int q = 0; q < rank; ++q) proj += g[i + (size_t)n * q] * solve[q];
      const double u = (col[i] - proj) / tau;
      g[i + (size_t)n * rank] = u;
      diag[i] -= u * u;
    }
    piv[rank] = pivot;
    ++rank;
    /* first maximum, plus the largest value strictly below it for the
     * decision margin (exact ties, e.g. residues still exactly 1 far from
     * every pivot, resolve to the first index on any implementation) */
    double best = diag[0];
    int bi = 0;
    for (int i = 1; i < n; ++i)
      if (diag[i] > best) {
        best = diag[i];
        bi = i;
      }
    double second = -INFINITY;
    for (int i = 0; i < n; ++i)
      if (diag[i] < best && diag[i] > second) second = diag[i];
    residue = best;
    pivot = bi;
    /* Decision margins, each relative to the mass a rounding difference can
     * move: a residue is 1 - sum_c g_c^2, so relative errors delta in the
     * terms move it by delta * (1 - residue).  Residues still ~1 far from
     * every pivot are therefore robust even when they sit a few ulps apart. */
    if (rank < maxr) {
      const double stop_gap = fabs(residue - tol) / fmax(1.0 - residue, tol);
      if (stop_gap < mg) mg = stop_gap;
      if (residue > tol) {
        const double mass = fmax((1.0 - best) + (1.0 - second), 2.220446049250313e-16);
        const double piv_gap = (best - second) / mass;
        if (piv_gap < mg) mg = piv_gap;
      }
    }
  }
  if (pivots) memcpy(pivots, piv, sizeof(int) * rank);
  if (margin) *margin = mg;
  free(diag); free(col); free(upper); free(solve); free(piv);
  return rank;
}

int orc_hsic_component(const double* x, int n, double sig, int maxr, double* gc, int* rank,
                       double* offdiag, double* moment, double* margin) {
  if (n < 6 || maxr <= 0 || !(sig > 0.0)) return -1;
  const int r = orc_inchol(x, n, 1.0 / sig, 0.001, maxr, gc, NULL, margin);
  double total = 0.0, dsq = 0.0;
  for (int c = 0; c < r; ++c) {
    double cs = 0.0;
    for (int i = 0; i < n; ++i) {
      const double v = gc[i + (size_t)n * c];
      cs += v;
      dsq += v * v;
    }
    total += cs * cs;
  }
  *offdiag = total - dsq;
  for (int c = 0; c < r; ++c) {
    double mean = 0.0;
    for (int i = 0; i < n; ++i) mean += gc[i + (size_t)n * c];
    mean /= (double)n;
    for (int i = 0; i < n; ++i) gc[i + (size_t)n * c] -= mean;
  }
  double m = 0.0;
  for (int a = 0; a < r; ++a)
    for (int b = 0; b < r; ++b) {
      double dot = 0.0;
      for (int i = 0; i < n; ++i) dot += gc[i + (size_t)n * a] * gc[i + (size_t)n * b];
      m += dot * dot;
    }
  *moment = m;
  *rank = r;
  return 0;
}

double orc_hsic_cross(const double* gx, int rx, const double* gy, int ry, int n,
                      const int* perm) {
  double s = 0.0;
  for (int a = 0; a < rx; ++a)
    for (int b = 0; b < ry; ++b) {
      double dot = 0.0;
      for (int i = 0; i < n; ++i) dot += gx[i + (size_t)n * a] * gy[perm[i] + (size_t)n * b];
      s += dot * dot;
    }
  return s;
}

void orc_hsic_gamma_from_factors(const double* gx, int rx, double offx, double momx,
                                 const double* gy, int ry, double offy, double momy, int n,
                                 double out[5]) {
  int* ident = (int*)malloc(sizeof(int) * n);
  for (int i = 0; i < n; ++i) ident[i] = i;
  const double nd = (double)n;
  const double hsic = orc_hsic_cross(gx, rx, gy, ry, n, ident) / (nd * nd);
  free(ident);
  const double mux = offx / (nd * (nd - 1.0));
  const double muy = offy / (nd * (nd - 1.0));
  const double mean = (1.0 + mux * muy - mux - muy) / nd;
  const double f = 2.0 * (nd - 4.0) * (nd - 5.0) / (nd * (nd - 1.0) * (nd - 2.0) * (nd - 3.0));
  const double var = f * momx * momy / (nd * nd * nd * nd);
  out[1] = hsic;
  out[2] = mean;
  out[3] = var;
  out[4] = 1e300;
  if (!isfinite(hsic) || !isfinite(mean) || !isfinite(var) || !(mean > 0.0) || !(var > 0.0))
    out[0] = 1.0;
  else
    out[0] = orc_gamma_pvalue(hsic, mean, var);
}
