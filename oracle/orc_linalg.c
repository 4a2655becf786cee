/*
 * Dense linear algebra for the CPU oracle (test infrastructure only).
 *
 * Plain C99, column-major storage, no BLAS/LAPACK dependency (none exists in
 * this image).  Each routine restates the semantics of the LAPACK/LINPACK call
 * the reference path relies on; operation order follows the textbook/LAPACK
 * algorithm so results agree with the reference at rounding level:
 *   - orc_pstf2_upper     : LAPACK dpstf2 'U' (pivoted Cholesky), used by mgcv
 *                           mroot and the magic aggregate root
 *                           (fastkpc/src/full_cuda_ci_native_setup.cpp:941-994,
 *                            fastkpc/src/mgcv_multi_penalty_cpp.cpp:133-170).
 *   - orc_geqp3           : LAPACK dgeqp3 (Householder QR, column pivoting)
 *                           (fastkpc/src/full_cuda_ci_native_setup.cpp:881-939).
 *   - orc_svd_jacobi      : thin SVD (one-sided Jacobi); stands in for dgesdd
 *                           (fastkpc/src/mgcv_multi_penalty_cpp.cpp:520-566).
 *   - orc_syev_jacobi     : symmetric eigen (cyclic Jacobi); stands in for dsyevr.
 *   - orc_tridiag_*       : symmetric tridiagonal reduction, implicit QL and
 *                           inverse iteration; used for Lanczos Ritz values
 *                           (dstedc in the reference) and dcov components
 *                           (eigs / eig_sym in the reference).
 */
#include "orc_internal.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define A_(i, j) a[(size_t)(j) * lda + (i)]

/* ---------------------------------------------------------------- pstf2 */
int orc_pstf2_upper(double* a, int n, int lda, int* piv, double tol, int* rank) {
  double* work = (double*)calloc((size_t)2 * n, sizeof(double));
  int info = 0;
  for (int i = 0; i < n; ++i) piv[i] = i;
  /* first pivot: first maximal diagonal */
  int pvt = 0;
  double ajj = A_(0, 0);
  for (int i = 1; i < n; ++i)
    if (A_(i, i) > ajj) { ajj = A_(i, i); pvt = i; }
  if (ajj <= 0.0 || isnan(ajj)) { *rank = 0; free(work); return 1; }
  double dstop = tol < 0.0 ? (double)n * DBL_EPSILON * ajj : tol;
  for (int i = 0; i < n; ++i) work[i] = 0.0;
  int j;
  for (j = 0; j < n; ++j) {
    for (int i = j; i < n; ++i) {
      if (j > 0) work[i] += A_(j - 1, i) * A_(j - 1, i);
      work[n + i] = A_(i, i) - work[i];
    }
    if (j > 0) {
      pvt = j;
      ajj = work[n + j];
      for (int i = j + 1; i < n; ++i)
        if (work[n + i] > ajj) { ajj = work[n + i]; pvt = i; }
      if (ajj <= dstop || isnan(ajj)) { A_(j, j) = ajj; info = 1; break; }
    }
    if (j != pvt) {
      A_(pvt, pvt) = A_(j, j);
      for (int k = 0; k < j; ++k) { double t = A_(k, j); A_(k, j) = A_(k, pvt); A_(k, pvt) = t; }
      for (int k = pvt + 1; k < n; ++k) { double t = A_(j, k); A_(j, k) = A_(pvt, k); A_(pvt, k) = t; }
      for (int k = j + 1; k < pvt; ++k) { double t = A_(j, k); A_(j, k) = A_(k, pvt); A_(k, pvt) = t; }
      double t = work[j]; work[j] = work[pvt]; work[pvt] = t;
      int it = piv[pvt]; piv[pvt] = piv[j]; piv[j] = it;
    }
    ajj = sqrt(ajj);
    A_(j, j) = ajj;
    if (j < n - 1) {
      for (int c = j + 1; c < n; ++c) {
        double s = 0.0;
        for (int k = 0; k < j; ++k) s += A_(k, c) * A_(k, j);
        A_(j, c) = A_(j, c) - s;
      }
      double r = 1.0 / ajj;
      for (int c = j + 1; c < n; ++c) A_(j, c) *= r;
    }
  }
  *rank = info ? j : n;
  free(work);
  return info;
}

/* ---------------------------------------------------------------- geqp3 */
static double col_norm(const double* x, int m) {
  double s = 0.0;
  for (int i = 0; i < m; ++i) s += x[i] * x[i];
  return sqrt(s);
}

/* Householder generator, LAPACK dlarfg convention: H = I - tau v v', v[0]=1.
 * x points to alpha followed by the m-1 trailing entries. */
void orc_larfg(int m, double* alpha, double* x, int incx, double* tau) {
  if (m <= 1) { *tau = 0.0; return; }
  double xn = 0.0;
  for (int i = 0; i < m - 1; ++i) xn += x[(size_t)i * incx] * x[(size_t)i * incx];
  xn = sqrt(xn);
  if (xn == 0.0) { *tau = 0.0; return; }
  double beta = -copysign(hypot(*alpha, xn), *alpha);
  *tau = (beta - *alpha) / beta;
  double sc = 1.0 / (*alpha - beta);
  for (int i = 0; i < m - 1; ++i) x[(size_t)i * incx] *= sc;
  *alpha = beta;
}

void orc_geqp3(double* a, int m, int n, int lda, int* jpvt, double* tau) {
  double* vn1 = (double*)malloc(sizeof(double) * n);
  double* vn2 = (double*)malloc(sizeof(double) * n);
  const double tol3z = sqrt(DBL_EPSILON);
  for (int j = 0; j < n; ++j) {
    jpvt[j] = j;
    vn1[j] = col_norm(&A_(0, j), m);
    vn2[j] = vn1[j];
  }
  int kmax = m < n ? m : n;
  for (int i = 0; i < kmax; ++i) {
    int pvt = i;
    for (int j = i + 1; j < n; ++j) if (vn1[j] > vn1[pvt]) pvt = j;
    if (pvt != i) {
      for (int r = 0; r < m; ++r) { double t = A_(r, pvt); A_(r, pvt) = A_(r, i); A_(r, i) = t; }
      int it = jpvt[pvt]; jpvt[pvt] = jpvt[i]; jpvt[i] = it;
      vn1[pvt] = vn1[i]; vn2[pvt] = vn2[i];
    }
    orc_larfg(m - i, &A_(i, i), &A_(i + 1, i), 1, &tau[i]);
    if (i < n - 1 && tau[i] != 0.0) {
      double aii = A_(i, i);
      A_(i, i) = 1.0;
      for (int j = i + 1; j < n; ++j) {
        double s = 0.0;
        for (int r = i; r < m; ++r) s += A_(r, i) * A_(r, j);
        s *= tau[i];
        for (int r = i; r < m; ++r) A_(r, j) -= s * A_(r, i);
      }
      A_(i, i) = aii;
    }
    for (int j = i + 1; j < n; ++j) {
      if (vn1[j] != 0.0) {
        double temp = fabs(A_(i, j)) / vn1[j];
        temp = 1.0 - temp * temp;
        if (temp < 0.0) temp = 0.0;
        double temp2 = temp * (vn1[j] / vn2[j]) * (vn1[j] / vn2[j]);
        if (temp2 <= tol3z) {
          vn1[j] = i + 1 < m ? col_norm(&A_(i + 1, j), m - i - 1) : 0.0;
          vn2[j] = vn1[j];
        } else {
          vn1[j] *= sqrt(temp);
        }
      }
    }
  }
  free(vn1);
  free(vn2);
}

/* y <- Q' y for the first k reflectors stored below the diagonal of a. */
void orc_apply_qt(const double* a, int m, int k, int lda, const double* tau, double* y) {
  for (int i = 0; i < k; ++i) {
    if (tau[i] == 0.0) continue;
    double s = y[i];
    for (int r = i + 1; r < m; ++r) s += A_(r, i) * y[r];
    s *= tau[i];
    y[i] -= s;
    for (int r = i + 1; r < m; ++r) y[r] -= s * A_(r, i);
  }
}

/* ---------------------------------------------------------------- SVD */
/* Thin SVD of m x n (m >= n) matrix a (column-major, lda=m) by one-sided
 * Jacobi.  On return u (m x n) holds left vectors, s (n) singular values in
 * decreasing order, v (n x n) right vectors. */
int orc_svd_jacobi(const double* a, int m, int n, double* u, double* s, double* v) {
  memcpy(u, a, sizeof(double) * (size_t)m * n);
  for (int i = 0; i < n * n; ++i) v[i] = 0.0;
  for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
  const double eps = DBL_EPSILON;
  int sweep;
  for (sweep = 0; sweep < 80; ++sweep) {
    int rotated = 0;
    for (int p = 0; p < n - 1; ++p) {
      for (int q = p + 1; q < n; ++q) {
        double* up = u + (size_t)p * m;
        double* uq = u + (size_t)q * m;
        double alpha = 0.0, beta = 0.0, gamma = 0.0;
        for (int i = 0; i < m; ++i) {
          alpha += up[i] * up[i];
          beta += uq[i] * uq[i];
          gamma += up[i] * uq[i];
        }
        if (gamma == 0.0 || fabs(gamma) <= eps * sqrt(alpha * beta)) continue;
        rotated = 1;
        double zeta = (beta - alpha) / (2.0 * gamma);
        double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        double c = 1.0 / sqrt(1.0 + t * t);
        double sn = c * t;
        for (int i = 0; i < m; ++i) {
          double x = up[i], y = uq[i];
          up[i] = c * x - sn * y;
          uq[i] = sn * x + c * y;
        }
        double* vp = v + (size_t)p * n;
        double* vq = v + (size_t)q * n;
        for (int i = 0; i < n; ++i) {
          double x = vp[i], y = vq[i];
          vp[i] = c * x - sn * y;
          vq[i] = sn * x + c * y;
        }
      }
    }
    if (!rotated) break;
  }
  for (int j = 0; j < n; ++j) {
    double nrm = col_norm(u + (size_t)j * m, m);
    s[j] = nrm;
    if (nrm > 0.0) {
      double r = 1.0 / nrm;
      for (int i = 0; i < m; ++i) u[(size_t)j * m + i] *= r;
    }
  }
  /* sort decreasing (selection sort keeps it simple) */
  for (int j = 0; j < n; ++j) {
    int best = j;
    for (int k = j + 1; k < n; ++k) if (s[k] > s[best]) best = k;
    if (best != j) {
      double t = s[j]; s[j] = s[best]; s[best] = t;
      for (int i = 0; i < m; ++i) { double x = u[(size_t)j * m + i]; u[(size_t)j * m + i] = u[(size_t)best * m + i]; u[(size_t)best * m + i] = x; }
      for (int i = 0; i < n; ++i) { double x = v[(size_t)j * n + i]; v[(size_t)j * n + i] = v[(size_t)best * n + i]; v[(size_t)best * n + i] = x; }
    }
  }
  return sweep >= 80 ? 1 : 0;
}

/* ---------------------------------------------------------------- syev */
/* Symmetric eigen by cyclic Jacobi; values ascending, vectors column-major.
 * A sweep is a deterministic function of (a, z): once a sweep changes no bit
 * of either, every later sweep is the identity, so the loop stops there with
 * exactly the state the full sweep budget would reach. */
static int bits_differ(double x, double y) {
  unsigned long long a, b;
  memcpy(&a, &x, sizeof a);
  memcpy(&b, &y, sizeof b);
  return a != b;
}

int orc_syev_jacobi(const double* a_in, int n, double* w, double* z) {
  double* a = (double*)malloc(sizeof(double) * (size_t)n * n);
  memcpy(a, a_in, sizeof(double) * (size_t)n * n);
  for (int i = 0; i < n * n; ++i) z[i] = 0.0;
  for (int i = 0; i < n; ++i) z[(size_t)i * n + i] = 1.0;
  int sweep;
  for (sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double x = a[(size_t)j * n + i] * a[(size_t)j * n + i];
        tot += x;
        if (i != j) off += x;
      }
    if (off <= 1e-300 || off <= DBL_EPSILON * DBL_EPSILON * tot * 1e-4) break;
    int changed = 0;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = a[(size_t)q * n + p];
        if (apq == 0.0) continue;
        double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = copysign(1.0, theta) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = a[(size_t)p * n + k], akq = a[(size_t)q * n + k];
          double np = c * akp - s * akq, nq = s * akp + c * akq;
          changed |= bits_differ(np, akp) | bits_differ(nq, akq);
          a[(size_t)p * n + k] = np;
          a[(size_t)q * n + k] = nq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = a[(size_t)k * n + p], aqk = a[(size_t)k * n + q];
          double np = c * apk - s * aqk, nq = s * apk + c * aqk;
          changed |= bits_differ(np, apk) | bits_differ(nq, aqk);
          a[(size_t)k * n + p] = np;
          a[(size_t)k * n + q] = nq;
        }
        for (int k = 0; k < n; ++k) {
          double zkp = z[(size_t)p * n + k], zkq = z[(size_t)q * n + k];
          double np = c * zkp - s * zkq, nq = s * zkp + c * zkq;
          changed |= bits_differ(np, zkp) | bits_differ(nq, zkq);
          z[(size_t)p * n + k] = np;
          z[(size_t)q * n + k] = nq;
        }
      }
    if (!changed) break;
  }
  for (int i = 0; i < n; ++i) w[i] = a[(size_t)i * n + i];
  for (int j = 0; j < n; ++j) {
    int best = j;
    for (int k = j + 1; k < n; ++k) if (w[k] < w[best]) best = k;
    if (best != j) {
      double t = w[j]; w[j] = w[best]; w[best] = t;
      for (int i = 0; i < n; ++i) { double x = z[(size_t)j * n + i]; z[(size_t)j * n + i] = z[(size_t)best * n + i]; z[(size_t)best * n + i] = x; }
    }
  }
  free(a);
  return sweep >= 100 ? 1 : 0;
}

/* ------------------------------------------------------- tridiagonal QL */
/* Implicit QL with Wilkinson shifts (EISPACK tql2/tqli).  d[0..n-1] diagonal,
 * e[0..n-2] off-diagonal (destroyed).  When z != NULL (n x n column-major,
 * initialised by the caller) eigenvectors are accumulated.  Eigenvalues are
 * left unsorted. */
int orc_tridiag_ql(double* d, double* e_in, int n, double* z, int ldz) {
  double* e = (double*)malloc(sizeof(double) * (size_t)(n + 1));
  for (int i = 0; i < n - 1; ++i) e[i] = e_in[i];
  e[n - 1] = 0.0;
  e[n] = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; ++m) {
        double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= DBL_EPSILON * dd) break;
      }
      if (m != l) {
        if (iter++ == 60) { free(e); return 1; }
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + copysign(r, g));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          e[i + 1] = (r = hypot(f, g));
          if (r == 0.0) { d[i + 1] -= p; e[m] = 0.0; break; }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          d[i + 1] = g + (p = s * r);
          g = c * r - b;
          if (z) {
            for (int k = 0; k < n; ++k) {
              f = z[(size_t)(i + 1) * ldz + k];
              z[(size_t)(i + 1) * ldz + k] = s * z[(size_t)i * ldz + k] + c * f;
              z[(size_t)i * ldz + k] = c * z[(size_t)i * ldz + k] - s * f;
            }
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  free(e);
  return 0;
}

/* Householder reduction of the symmetric n x n matrix a (full storage,
 * column-major) to tridiagonal form Q' A Q = T.  Reflector k (k=0..n-3) acts
 * on rows/cols k+1..n-1 with vector v (v[0]=1) stored in a(k+2.., k) and
 * scalar tau[k].  d (n) and e (n-1) receive T. */
void orc_tridiag_reduce(double* a, int n, int lda, double* d, double* e, double* tau) {
  double* p = (double*)malloc(sizeof(double) * (size_t)n);
  for (int k = 0; k < n - 2; ++k) {
    int m = n - k - 1; /* length of the reflector */
    double alpha = A_(k + 1, k);
    orc_larfg(m, &alpha, &A_(k + 2, k), 1, &tau[k]);
    e[k] = alpha;
    double tk = tau[k];
    if (tk != 0.0) {
      /* v = [1; a(k+2:n-1,k)] on indices k+1..n-1 */
      A_(k + 1, k) = 1.0;
      /* p = tau * A22 * v */
      for (int i = k + 1; i < n; ++i) {
        double s = 0.0;
        for (int j = k + 1; j < n; ++j) s += A_(i, j) * A_(j, k);
        p[i] = tk * s;
      }
      double pv = 0.0;
      for (int i = k + 1; i < n; ++i) pv += p[i] * A_(i, k);
      double half = -0.5 * tk * pv;
      for (int i = k + 1; i < n; ++i) p[i] += half * A_(i, k); /* w */
      for (int j = k + 1; j < n; ++j)
        for (int i = k + 1; i < n; ++i)
          A_(i, j) -= A_(i, k) * p[j] + p[i] * A_(j, k);
    }
    A_(k + 1, k) = e[k];
    d[k] = A_(k, k);
  }
  if (n >= 2) {
    d[n - 2] = A_(n - 2, n - 2);
    e[n - 2] = A_(n - 1, n - 2);
  }
  d[n - 1] = A_(n - 1, n - 1);
  if (n >= 2) tau[n - 2] = 0.0;
  free(p);
}

/* Apply Q = H_0 H_1 ... H_{n-3} (from orc_tridiag_reduce) to the n x k block
 * zz (column-major, ldz = n): zz <- Q zz. */
void orc_tridiag_backtransform(const double* a, int n, int lda, const double* tau,
                               double* zz, int k) {
  for (int r = n - 3; r >= 0; --r) {
    double t = tau[r];
    if (t == 0.0) continue;
    for (int c = 0; c < k; ++c) {
      double* z = zz + (size_t)c * n;
      double s = z[r + 1];
      for (int i = r + 2; i < n; ++i) s += A_(i, r) * z[i];
      s *= t;
      z[r + 1] -= s;
      for (int i = r + 2; i < n; ++i) z[i] -= s * A_(i, r);
    }
  }
}

/* Eigenvector of the tridiagonal (d, e) for eigenvalue lam by inverse
 * iteration (LU with partial pivoting, as in LAPACK dstein), orthogonalised
 * against the `nprev` columns in prev (already computed vectors whose
 * eigenvalues lie within the cluster tolerance).  Result normalised. */
void orc_tridiag_invit(const double* d, const double* e, int n, double lam,
                       const double* prev, int nprev, double* z, unsigned* seed) {
  double* dl = (double*)malloc(sizeof(double) * (size_t)n * 5);
  double* dm = dl + n;
  double* du = dm + n;
  double* du2 = du + n;
  int* ip = (int*)malloc(sizeof(int) * (size_t)n);
  double tnorm = 0.0;
  for (int i = 0; i < n; ++i) {
    double s = fabs(d[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
    if (s > tnorm) tnorm = s;
  }
  double pert = DBL_EPSILON * tnorm;
  /* factor T - lam I = L U with partial pivoting (dlagtf style) */
  for (int i = 0; i < n; ++i) dm[i] = d[i] - lam;
  for (int i = 0; i < n - 1; ++i) { dl[i] = e[i]; du[i] = e[i]; }
  for (int i = 0; i < n; ++i) du2[i] = 0.0;
  for (int i = 0; i < n - 1; ++i) {
    if (fabs(dm[i]) >= fabs(dl[i])) {
      ip[i] = 0;
      if (dm[i] == 0.0) dm[i] = pert;
      double f = dl[i] / dm[i];
      dl[i] = f;
      dm[i + 1] -= f * du[i];
    } else {
      ip[i] = 1;
      double f = dm[i] / dl[i];
      dm[i] = dl[i];
      dl[i] = f;
      double t = du[i];
      du[i] = dm[i + 1];
      dm[i + 1] = t - f * dm[i + 1];
      if (i < n - 2) { du2[i] = du[i + 1]; du[i + 1] = -f * du[i + 1]; }
    }
  }
  if (dm[n - 1] == 0.0) dm[n - 1] = pert;
  for (int i = 0; i < n; ++i) {
    *seed = *seed * 1103515245u + 12345u;
    z[i] = ((double)((*seed >> 8) & 0xFFFF) / 65536.0) - 0.5;
  }
  for (int it = 0; it < 5; ++it) {
    /* orthogonalise against cluster members */
    for (int c = 0; c < nprev; ++c) {
      const double* q = prev + (size_t)c * n;
      double s = 0.0;
      for (int i = 0; i < n; ++i) s += q[i] * z[i];
      for (int i = 0; i < n; ++i) z[i] -= s * q[i];
    }
    /* solve L U x = z */
    for (int i = 0; i < n - 1; ++i) {
      if (ip[i] == 0) z[i + 1] -= dl[i] * z[i];
      else { double t = z[i]; z[i] = z[i + 1]; z[i + 1] = t - dl[i] * z[i]; }
    }
    z[n - 1] /= dm[n - 1];
    if (n >= 2) z[n - 2] = (z[n - 2] - du[n - 2] * z[n - 1]) / dm[n - 2];
    for (int i = n - 3; i >= 0; --i) z[i] = (z[i] - du[i] * z[i + 1] - du2[i] * z[i + 2]) / dm[i];
    double nrm = 0.0;
    for (int i = 0; i < n; ++i) nrm += z[i] * z[i];
    nrm = sqrt(nrm);
    for (int i = 0; i < n; ++i) z[i] /= nrm;
  }
  for (int c = 0; c < nprev; ++c) {
    const double* q = prev + (size_t)c * n;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += q[i] * z[i];
    for (int i = 0; i < n; ++i) z[i] -= s * q[i];
  }
  double nrm = 0.0;
  for (int i = 0; i < n; ++i) nrm += z[i] * z[i];
  nrm = sqrt(nrm);
  for (int i = 0; i < n; ++i) z[i] /= nrm;
  free(dl);
  free(ip);
}
