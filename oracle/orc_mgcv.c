/*
 * regrXonS for the CPU oracle (test infrastructure only): mgcv 1.9-1 `gam`
 * with thin-plate regression smooths and GCV.Cp smoothing selection via
 * `magic`, restated in plain C.
 *
 * Follows (reference paths):
 *   kpcalg/R/regrXonS.R:45-57           formula choice: s(x1,..) for |S|<=2,
 *                                       additive s(x1)+...+s(xk) for |S|>2
 *   fastkpc/src/full_cuda_ci_native_setup.cpp
 *     :96-107   R mean for the smooth shift
 *     :114-152  thin-plate eta constant and eta(r^2)
 *     :154-198  null-space dimension and polynomial powers
 *     :200-239  sorted unique locations
 *     :303-494  Rlanczos (LCG start vector, 2-pass reorthogonalisation,
 *               check cadence, eps^0.7 stop rule incl. the lower-end index
 *               quirk at :439)
 *     :496-594  QT Householders, hq_post, hq_pre_transpose
 *     :621-688  sum-to-zero constraint via LINPACK dqrdc2/dqrqty
 *     :690-858  build_smooth (X=[UDZ|T], penalty, RMS scaling, S scaling)
 *     :1188-1317, :1321-1356 assembly X=[1|X_1|...], penalty offsets
 *     :1019-1127 initial.sp
 *   fastkpc/src/mgcv_multi_penalty_cpp.cpp
 *     :612-700  problem preparation (pivoted QR, pivoted penalty roots)
 *     :721-751  aggregate penalty and its pivoted Cholesky root
 *     :752-818  GCV core: SVD of [R; root], rank tol, rss, edf, score
 *     :819-990  gradient/Hessian of the GCV score in log(sp)
 *     :1118-1346 Newton/steepest descent with step halving, convergence
 *               test, final boundary ("infinity") probes
 *   fastkpc/src/cuda/mgcv_single_penalty_gcv.cu:29-39 single-penalty
 *     control constants (15 halvings).
 */
#include "kpc_oracle.h"
#include "orc_internal.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define M_(a, ld, i, j) (a)[(size_t)(j) * (ld) + (i)]

/* ------------------------------------------------------- thin plate */
static double eta_constant(int m, int d) {
  const double pi = asin(1.0) * 2.0;
  const double gamma_half = sqrt(pi);
  const int d2 = d / 2;
  const int m2 = 2 * m;
  double value = 1.0;
  if (d % 2 == 0) {
    value = ((m + 1 + d2) % 2) ? -1.0 : 1.0;
    for (int i = 0; i < m2 - 1; ++i) value /= 2.0;
    for (int i = 0; i < d2; ++i) value /= pi;
    for (int i = 2; i < m; ++i) value /= (double)i;
    for (int i = 2; i <= m - d2; ++i) value /= (double)i;
  } else {
    value = gamma_half;
    const int k = m - (d - 1) / 2;
    for (int i = 0; i < k; ++i) value /= -0.5 - (double)i;
    for (int i = 0; i < m; ++i) value /= 4.0;
    for (int i = 0; i < d2; ++i) value /= pi;
    value /= gamma_half;
    for (int i = 2; i < m; ++i) value /= (double)i;
  }
  return value;
}

static double eta_r2(int m, int d, double r2, double eta0) {
  if (r2 <= 0.0) return 0.0;
  const int d2 = d / 2;
  double value = eta0;
  if (d % 2 == 0) {
    value *= log(r2) * 0.5;
    for (int i = 0; i < m - d2; ++i) value *= r2;
  } else {
    for (int i = 0; i < m - d2 - 1; ++i) value *= r2;
    value *= sqrt(r2);
  }
  return value;
}

static int null_dim(int d, int m) {
  int dim = 1;
  for (int i = 0; i < d; ++i) dim *= d + m - 1 - i;
  for (int i = 2; i <= d; ++i) dim /= i;
  return dim;
}

/* powers[term + M*coord] */
static void poly_powers(int d, int order, int M, int* powers) {
  int index[4] = {0, 0, 0, 0};
  for (int term = 0; term < M; ++term) {
    for (int c = 0; c < d; ++c) powers[term + M * c] = index[c];
    int sum = 0;
    for (int c = 0; c < d; ++c) sum += index[c];
    if (sum < order - 1) {
      ++index[0];
    } else {
      sum -= index[0];
      index[0] = 0;
      for (int c = 1; c < d; ++c) {
        ++index[c];
        ++sum;
        if (sum == order) {
          sum -= index[c];
          index[c] = 0;
        } else {
          break;
        }
      }
    }
  }
}

/* Symmetric tridiagonal eigen with vectors, sorted descending (the reference
 * reverses LAPACK's ascending dstedc output). z is size x size. */
static void tri_eigen_desc(double* d, const double* e, int size, double* z) {
  double* ee = (double*)malloc(sizeof(double) * (size_t)(size + 1));
  for (int i = 0; i < size - 1; ++i) ee[i] = e[i];
  for (int i = 0; i < size * size; ++i) z[i] = 0.0;
  for (int i = 0; i < size; ++i) z[(size_t)i * size + i] = 1.0;
  orc_tridiag_ql(d, ee, size, z, size);
  /* sort descending */
  for (int j = 0; j < size; ++j) {
    int best = j;
    for (int k = j + 1; k < size; ++k) if (d[k] > d[best]) best = k;
    if (best != j) {
      double t = d[j]; d[j] = d[best]; d[best] = t;
      for (int i = 0; i < size; ++i) {
        double x = z[(size_t)j * size + i];
        z[(size_t)j * size + i] = z[(size_t)best * size + i];
        z[(size_t)best * size + i] = x;
      }
    }
  }
  free(ee);
}

/* mgcv Rlanczos restatement: largest-magnitude eigenpairs of symmetric E. */
static int lanczos_lm(const double* E, int n, int k, double* vecs, double* vals) {
  const double tol = pow(DBL_EPSILON, 0.7);
  int upper_req = k, lower_req = 0;
  int check = k / 2;
  if (check < 10) check = 10;
  int tenth = (int)floor(n / 10.0);
  if (tenth < 1) tenth = 1;
  if (tenth < check) check = tenth;
  double* q = (double*)calloc((size_t)n * (n + 1), sizeof(double));
  unsigned long st = 1;
  double nrm = 0.0;
  for (int r = 0; r < n; ++r) {
    st = (st * 106UL + 1283UL) % 6075UL;
    double v = (double)st / 6075.0 - 0.5;
    q[r] = v;
    nrm += v * v;
  }
  nrm = sqrt(nrm);
  for (int r = 0; r < n; ++r) q[r] /= nrm;
  double* alpha = (double*)calloc(n, sizeof(double));
  double* beta = (double*)calloc(n, sizeof(double));
  double* z = (double*)calloc(n, sizeof(double));
  double* errors = (double*)malloc(sizeof(double) * n);
  double* tv = NULL;
  double* tvals = (double*)calloc(n, sizeof(double));
  int tri_len = 0, completed = n;
  for (int it = 0; it < n; ++it) {
    const double* qi = q + (size_t)it * n;
    /* dsymv with the upper triangle */
    for (int r = 0; r < n; ++r) z[r] = 0.0;
    for (int c = 0; c < n; ++c) {
      double t1 = qi[c], t2 = 0.0;
      for (int r = 0; r < c; ++r) {
        z[r] += t1 * E[(size_t)c * n + r];
        t2 += E[(size_t)c * n + r] * qi[r];
      }
      z[c] += t1 * E[(size_t)c * n + c] + t2;
    }
    double proj = 0.0;
    for (int r = 0; r < n; ++r) proj += qi[r] * z[r];
    alpha[it] = proj;
    if (it == 0) {
      for (int r = 0; r < n; ++r) z[r] -= proj * q[r];
    } else {
      const double prior = beta[it - 1];
      const double* qp = q + (size_t)(it - 1) * n;
      for (int r = 0; r < n; ++r) z[r] -= proj * qi[r] + prior * qp[r];
      for (int pass = 0; pass < 2; ++pass) {
        for (int b = 0; b <= it; ++b) {
          const double* qb = q + (size_t)b * n;
          double c = 0.0;
          for (int r = 0; r < n; ++r) c += z[r] * qb[r];
          c = -c;
          for (int r = 0; r < n; ++r) z[r] += c * qb[r];
        }
      }
    }
    double rn = 0.0;
    for (int r = 0; r < n; ++r) rn += z[r] * z[r];
    rn = sqrt(rn);
    beta[it] = rn;
    if (it < n - 1) {
      double* qn = q + (size_t)(it + 1) * n;
      for (int r = 0; r < n; ++r) qn[r] = z[r] / rn;
    }
    if ((it >= k && it % check == 0) || it == n - 1) {
      const int size = it + 1;
      for (int i = 0; i < size; ++i) tvals[i] = alpha[i];
      free(tv);
      tv = (double*)malloc(sizeof(double) * (size_t)size * size);
      tri_eigen_desc(tvals, beta, size, tv);
      tri_len = size;
      double tnorm = fabs(tvals[0]);
      if (fabs(tvals[it]) > tnorm) tnorm = fabs(tvals[it]);
      for (int e = 0; e < size; ++e) errors[e] = fabs(beta[it] * tv[(size_t)e * size + it]);
      if (it >= k) {
        const double maxerr = tnorm * tol;
        int pos = 0, neg = 0, conv = 1;
        while (pos + neg < k) {
          if (fabs(tvals[pos]) >= fabs(tvals[it - neg])) {
            if (errors[pos] > maxerr) { conv = 0; break; }
            ++pos;
          } else {
            if (errors[neg] > maxerr) { conv = 0; break; } /* reference index quirk */
            ++neg;
          }
        }
        if (conv) {
          upper_req = pos;
          lower_req = neg;
          completed = it + 1;
          break;
        }
      }
    }
  }
  for (int i = 0; i < n * k; ++i) vecs[i] = 0.0;
  for (int e = 0; e < upper_req + lower_req; ++e) {
    int src = e < upper_req ? e : completed - (upper_req + lower_req - e);
    vals[e] = tvals[src];
    for (int b = 0; b < completed; ++b) {
      double coef = tv[(size_t)src * tri_len + b];
      const double* qb = q + (size_t)b * n;
      for (int r = 0; r < n; ++r) vecs[(size_t)e * n + r] += qb[r] * coef;
    }
  }
  free(q); free(alpha); free(beta); free(z); free(errors); free(tv); free(tvals);
  return completed;
}

typedef struct {
  int n, cols;      /* X is n x cols (after constraint: basis-1) */
  double* X;        /* column-major */
  double* S;        /* cols x cols column-major */
  int rank;
} smooth_t;

static void free_smooth(smooth_t* s) { free(s->X); free(s->S); }

/* Row-major helpers operate on (rows x cols) stored row-major, matching the
 * reference's RowMatrix for the Householder pieces. */
static void qt_householders(double* storage, double* in, int rows, int cols) {
  for (int step = 0; step < rows; ++step) {
    double* rs = in + (size_t)step * cols;
    double scale = 0.0;
    for (int c = 0; c < cols - step; ++c) if (fabs(rs[c]) > scale) scale = fabs(rs[c]);
    if (scale != 0.0) for (int c = 0; c < cols - step; ++c) rs[c] /= scale;
    double len = 0.0;
    for (int c = 0; c < cols - step; ++c) len += rs[c] * rs[c];
    len = sqrt(len);
    if (rs[cols - step - 1] < 0.0) len = -len;
    rs[cols - step - 1] += len;
    double mult = len != 0.0 ? 1.0 / (len * rs[cols - step - 1]) : 0.0;
    double terminal = len * scale;
    for (int r = step + 1; r < rows; ++r) {
      double* rr = in + (size_t)r * cols;
      double pr = 0.0;
      for (int c = 0; c < cols - step; ++c) pr += rs[c] * rr[c];
      pr *= mult;
      for (int c = 0; c < cols - step; ++c) rr[c] -= pr * rs[c];
    }
    double rm = sqrt(mult);
    for (int c = 0; c < cols - step; ++c) storage[(size_t)step * cols + c] = rs[c] * rm;
    for (int c = cols - step; c < cols; ++c) storage[(size_t)step * cols + c] = 0.0;
    rs[cols - step - 1] = -terminal;
    for (int c = 0; c < cols - step - 1; ++c) rs[c] = 0.0;
  }
}

/* m (rows x cols, row-major) <- m * prod(I - h h') */
static void hq_post(double* m, int rows, int cols, const double* h, int nh) {
  double* prod = (double*)malloc(sizeof(double) * rows);
  for (int t = 0; t < nh; ++t) {
    const double* ht = h + (size_t)t * cols;
    for (int r = 0; r < rows; ++r) {
      double v = 0.0;
      for (int c = 0; c < cols; ++c) v += m[(size_t)r * cols + c] * ht[c];
      prod[r] = v;
    }
    for (int r = 0; r < rows; ++r)
      for (int c = 0; c < cols; ++c) m[(size_t)r * cols + c] -= prod[r] * ht[c];
  }
  free(prod);
}

static void hq_pre_t(double* m, int rows, int cols, const double* h, int nh) {
  double* prod = (double*)malloc(sizeof(double) * cols);
  for (int t = 0; t < nh; ++t) {
    const double* ht = h + (size_t)t * rows;
    for (int c = 0; c < cols; ++c) {
      double v = 0.0;
      for (int r = 0; r < rows; ++r) v += m[(size_t)r * cols + c] * ht[r];
      prod[c] = v;
    }
    for (int r = 0; r < rows; ++r)
      for (int c = 0; c < cols; ++c) m[(size_t)r * cols + c] -= prod[c] * ht[r];
  }
  free(prod);
}

/* LINPACK dqrdc2 on a single column + dqrqty, i.e. one Householder
 * reflection in LINPACK's convention. c (len k) is overwritten by the packed
 * QR; *qraux receives the auxiliary scalar. */
static void linpack_qr1(double* c, int k, double* qraux) {
  double nrm = 0.0;
  for (int i = 0; i < k; ++i) nrm += c[i] * c[i];
  nrm = sqrt(nrm);
  *qraux = 0.0;
  if (k == 1 || nrm == 0.0) return;
  if (c[0] != 0.0) nrm = copysign(nrm, c[0]);
  double r = 1.0 / nrm;
  for (int i = 0; i < k; ++i) c[i] *= r;
  c[0] = 1.0 + c[0];
  *qraux = c[0];
  c[0] = -nrm;
}

static void linpack_qty1(const double* c, int k, double qraux, double* y) {
  if (qraux == 0.0) return;
  double t = qraux * y[0];
  for (int i = 1; i < k; ++i) t += c[i] * y[i];
  t = -t / qraux;
  y[0] += t * qraux;
  for (int i = 1; i < k; ++i) y[i] += t * c[i];
}

/* cond: n x d column-major slice of the data (d = 1 or 2 columns). */
static int build_smooth(const double* data, int n, int ld, const int* cols, int d,
                        int kb, smooth_t* out) {
  int order = 0;
  while (2 * order < d + 2) ++order;
  const int M = null_dim(d, order);
  double* sh = (double*)malloc(sizeof(double) * (size_t)n * d); /* row-major n x d */
  for (int c = 0; c < d; ++c) {
    double mu = orc_r_mean(data + (size_t)cols[c] * ld, n, 1);
    for (int r = 0; r < n; ++r) sh[(size_t)r * d + c] = data[(size_t)cols[c] * ld + r] - mu;
  }
  /* sorted unique locations (lexicographic, stable by row index) */
  int* ord = (int*)malloc(sizeof(int) * n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  for (int i = 1; i < n; ++i) {
    int oi = ord[i], j = i - 1;
    while (j >= 0) {
      int oj = ord[j], less = 0, decided = 0;
      for (int c = 0; c < d && !decided; ++c) {
        double a = sh[(size_t)oi * d + c], b = sh[(size_t)oj * d + c];
        if (a < b) { less = 1; decided = 1; } else if (a > b) { decided = 1; }
      }
      if (!decided) less = oi < oj;
      if (!less) break;
      ord[j + 1] = ord[j];
      --j;
    }
    ord[j + 1] = oi;
  }
  int* map = (int*)malloc(sizeof(int) * n);
  int* uo = (int*)malloc(sizeof(int) * n);
  int nu = 0;
  for (int t = 0; t < n; ++t) {
    int idx = ord[t], is_new = nu == 0;
    if (!is_new) {
      int prior = uo[nu - 1];
      for (int c = 0; c < d; ++c)
        if (sh[(size_t)idx * d + c] != sh[(size_t)prior * d + c]) { is_new = 1; break; }
    }
    if (is_new) uo[nu++] = idx;
    map[idx] = nu - 1;
  }
  if (nu < kb || nu > 2000) { free(sh); free(ord); free(map); free(uo); return -3; }
  double* U = (double*)malloc(sizeof(double) * (size_t)nu * d);
  for (int r = 0; r < nu; ++r)
    for (int c = 0; c < d; ++c) U[(size_t)r * d + c] = sh[(size_t)uo[r] * d + c];
  const double eta0 = eta_constant(order, d);
  double* E = (double*)calloc((size_t)nu * nu, sizeof(double));
  for (int r = 0; r < nu; ++r)
    for (int p = 0; p < r; ++p) {
      double s2 = 0.0;
      for (int c = 0; c < d; ++c) {
        double df = U[(size_t)r * d + c] - U[(size_t)p * d + c];
        s2 += df * df;
      }
      double v = eta_r2(order, d, s2, eta0);
      E[(size_t)p * nu + r] = v;
      E[(size_t)r * nu + p] = v;
    }
  int powers[16];
  poly_powers(d, order, M, powers);
  double* T = (double*)malloc(sizeof(double) * (size_t)nu * M); /* row-major nu x M */
  for (int r = 0; r < nu; ++r)
    for (int t = 0; t < M; ++t) {
      double v = 1.0;
      for (int c = 0; c < d; ++c)
        for (int e = 0; e < powers[t + M * c]; ++e) v *= U[(size_t)r * d + c];
      T[(size_t)r * M + t] = v;
    }
  double* ev = (double*)malloc(sizeof(double) * (size_t)nu * kb);
  double* eval = (double*)malloc(sizeof(double) * kb);
  lanczos_lm(E, nu, kb, ev, eval);
  /* TU = T' U  (M x kb, row-major) */
  double* TU = (double*)calloc((size_t)M * kb, sizeof(double));
  for (int sr = 0; sr < nu; ++sr)
    for (int orow = 0; orow < M; ++orow) {
      double mult = T[(size_t)sr * M + orow];
      for (int c = 0; c < kb; ++c) TU[(size_t)orow * kb + c] += mult * ev[(size_t)c * nu + sr];
    }
  double* Z = (double*)calloc((size_t)M * kb, sizeof(double));
  qt_householders(Z, TU, M, kb);
  double* Xu = (double*)malloc(sizeof(double) * (size_t)nu * kb); /* row-major */
  for (int r = 0; r < nu; ++r)
    for (int c = 0; c < kb; ++c) Xu[(size_t)r * kb + c] = ev[(size_t)c * nu + r] * eval[c];
  hq_post(Xu, nu, kb, Z, M);
  for (int r = 0; r < nu; ++r) {
    for (int c = kb - M; c < kb; ++c) Xu[(size_t)r * kb + c] = 0.0;
    for (int t = 0; t < M; ++t) Xu[(size_t)r * kb + kb - M + t] = T[(size_t)r * M + t];
  }
  double* X = (double*)malloc(sizeof(double) * (size_t)n * kb); /* row-major */
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < kb; ++c) X[(size_t)r * kb + c] = Xu[(size_t)map[r] * kb + c];
  double* P = (double*)calloc((size_t)kb * kb, sizeof(double)); /* row-major */
  for (int i = 0; i < kb; ++i) P[(size_t)i * kb + i] = eval[i];
  hq_post(P, kb, kb, Z, M);
  hq_pre_t(P, kb, kb, Z, M);
  for (int r = 0; r < kb; ++r)
    for (int c = kb - M; c < kb; ++c) { P[(size_t)r * kb + c] = 0.0; P[(size_t)c * kb + r] = 0.0; }
  for (int c = 0; c < kb; ++c) {
    double ss = 0.0;
    for (int r = 0; r < n; ++r) ss += X[(size_t)r * kb + c] * X[(size_t)r * kb + c];
    double rms = sqrt(ss / (double)n);
    for (int r = 0; r < n; ++r) X[(size_t)r * kb + c] /= rms;
    for (int i = 0; i < kb; ++i) P[(size_t)c * kb + i] /= rms;
    for (int i = 0; i < kb; ++i) P[(size_t)i * kb + c] /= rms;
  }
  double* SP = (double*)malloc(sizeof(double) * (size_t)kb * kb);
  for (int r = 0; r < kb; ++r)
    for (int c = 0; c < kb; ++c) SP[(size_t)r * kb + c] = (P[(size_t)r * kb + c] + P[(size_t)c * kb + r]) / 2.0;
  double xnorm = 0.0; /* max abs row sum */
  for (int r = 0; r < n; ++r) {
    double s = 0.0;
    for (int c = 0; c < kb; ++c) s += fabs(X[(size_t)r * kb + c]);
    if (s > xnorm) xnorm = s;
  }
  double pnorm = 0.0; /* max abs column sum */
  for (int c = 0; c < kb; ++c) {
    double s = 0.0;
    for (int r = 0; r < kb; ++r) s += fabs(SP[(size_t)r * kb + c]);
    if (s > pnorm) pnorm = s;
  }
  double pscale = pnorm / (xnorm * xnorm);
  for (int i = 0; i < kb * kb; ++i) SP[i] /= pscale;
  /* centring constraint absorption */
  double* cm = (double*)malloc(sizeof(double) * kb);
  for (int c = 0; c < kb; ++c) {
    long double t = 0.0L;
    for (int r = 0; r < n; ++r) t += X[(size_t)r * kb + c];
    cm[c] = (double)(t / (long double)n);
  }
  double qraux;
  linpack_qr1(cm, kb, &qraux);
  out->n = n;
  out->cols = kb - 1;
  out->rank = kb - M;
  out->X = (double*)malloc(sizeof(double) * (size_t)n * (kb - 1));
  double* yv = (double*)malloc(sizeof(double) * kb);
  for (int r = 0; r < n; ++r) {
    for (int c = 0; c < kb; ++c) yv[c] = X[(size_t)r * kb + c];
    linpack_qty1(cm, kb, qraux, yv);
    for (int c = 0; c < kb - 1; ++c) out->X[(size_t)c * n + r] = yv[c + 1];
  }
  /* left = Q' S (columns of S), then right = Q' (left rows 1..)' */
  double* left = (double*)malloc(sizeof(double) * (size_t)kb * kb); /* column-major */
  for (int c = 0; c < kb; ++c) {
    for (int r = 0; r < kb; ++r) yv[r] = SP[(size_t)r * kb + c];
    linpack_qty1(cm, kb, qraux, yv);
    for (int r = 0; r < kb; ++r) left[(size_t)c * kb + r] = yv[r];
  }
  out->S = (double*)malloc(sizeof(double) * (size_t)(kb - 1) * (kb - 1));
  for (int fr = 0; fr < kb - 1; ++fr) {
    for (int oc = 0; oc < kb; ++oc) yv[oc] = left[(size_t)oc * kb + fr + 1];
    linpack_qty1(cm, kb, qraux, yv);
    for (int c = 0; c < kb - 1; ++c) out->S[(size_t)c * (kb - 1) + fr] = yv[c + 1];
  }
  free(sh); free(ord); free(map); free(uo); free(U); free(E); free(T); free(ev); free(eval);
  free(TU); free(Z); free(Xu); free(X); free(P); free(SP); free(cm); free(yv); free(left);
  return 0;
}

/* ------------------------------------------------------------- setup */
void orc_setup_free(orc_setup* s) {
  free(s->X);
  for (int i = 0; i < s->m; ++i) free(s->S[i]);
  free(s->S);
  free(s->off);
  free(s->dim);
  free(s->rank);
  memset(s, 0, sizeof(*s));
}

int orc_regr_setup(const double* data, int n, int ld, const int* S, int s_len, orc_setup* out) {
  memset(out, 0, sizeof(*out));
  if (s_len < 1) return -1;
  int nsm = s_len <= 2 ? 1 : s_len;
  smooth_t* sm = (smooth_t*)calloc(nsm, sizeof(smooth_t));
  int rc = 0;
  if (s_len <= 2) {
    rc = build_smooth(data, n, ld, S, s_len, s_len == 1 ? 10 : 30, &sm[0]);
  } else {
    for (int i = 0; i < s_len && rc == 0; ++i) rc = build_smooth(data, n, ld, &S[i], 1, 10, &sm[i]);
  }
  if (rc != 0) { for (int i = 0; i < nsm; ++i) free_smooth(&sm[i]); free(sm); return rc; }
  int q = 1;
  for (int i = 0; i < nsm; ++i) q += sm[i].cols;
  if (q >= n) { for (int i = 0; i < nsm; ++i) free_smooth(&sm[i]); free(sm); return -4; }
  out->n = n;
  out->q = q;
  out->m = nsm;
  out->X = (double*)malloc(sizeof(double) * (size_t)n * q);
  out->S = (double**)calloc(nsm, sizeof(double*));
  out->off = (int*)malloc(sizeof(int) * nsm);
  out->dim = (int*)malloc(sizeof(int) * nsm);
  out->rank = (int*)malloc(sizeof(int) * nsm);
  for (int r = 0; r < n; ++r) out->X[r] = 1.0;
  int dest = 1;
  for (int i = 0; i < nsm; ++i) {
    for (int c = 0; c < sm[i].cols; ++c)
      memcpy(out->X + (size_t)(dest + c) * n, sm[i].X + (size_t)c * n, sizeof(double) * n);
    out->S[i] = sm[i].S;
    sm[i].S = NULL;
    out->off[i] = dest;
    out->dim[i] = sm[i].cols;
    out->rank[i] = sm[i].rank;
    dest += sm[i].cols;
    free_smooth(&sm[i]);
  }
  free(sm);
  return 0;
}

/* -------------------------------------------------------------- magic */
typedef struct {
  int n, q, m;
  double rank_tol;
  const double* X;
  double* R;        /* q x q */
  double* qr;       /* packed QR of X (n x q) */
  double* tau;
  int* piv;         /* pivot[i] = original column of pivoted column i */
  double** rS;      /* m roots, q x rank_i (pivoted rows) */
  int* rrank;
  double** Sp;      /* m pivoted penalties rS rS' (q x q) */
  double* init_lsp;
} magic_problem;

typedef struct {
  int nr, arank;
  double rss, edf, score, cond;
  double rank_margin;  /* min_i |d_i/d_0 - rank_tol| / rank_tol */
  double *U1, *V, *d, *y1; /* U1 q x nr, V q x nr */
} magic_core;

static void core_free(magic_core* c) { free(c->U1); free(c->V); free(c->d); free(c->y1); memset(c, 0, sizeof(*c)); }

/* mroot of a penalty block (dpstrf tol 0): returns dim x rank root (column-major) */
static double* block_root(const double* S, int dim, int rank) {
  double* f = (double*)malloc(sizeof(double) * (size_t)dim * dim);
  memcpy(f, S, sizeof(double) * (size_t)dim * dim);
  for (int c = 0; c < dim; ++c)
    for (int r = c + 1; r < dim; ++r) f[(size_t)c * dim + r] = 0.0;
  int* piv = (int*)malloc(sizeof(int) * dim);
  int ar = 0;
  orc_pstf2_upper(f, dim, dim, piv, 0.0, &ar);
  if (ar < rank) { free(f); free(piv); return NULL; }
  if (ar < dim)
    for (int c = ar; c < dim; ++c)
      for (int r = ar; r < dim; ++r) f[(size_t)c * dim + r] = 0.0;
  int* inv = (int*)malloc(sizeof(int) * dim);
  for (int p = 0; p < dim; ++p) inv[piv[p]] = p;
  double* root = (double*)malloc(sizeof(double) * (size_t)dim * rank);
  for (int o = 0; o < dim; ++o) {
    int rc = inv[o];
    for (int c = 0; c < rank; ++c) root[(size_t)c * dim + o] = f[(size_t)rc * dim + c];
  }
  free(f); free(piv); free(inv);
  return root;
}

static void initial_lsp(const orc_setup* s, double* out) {
  int q = s->q;
  double* xd = (double*)malloc(sizeof(double) * q);
  for (int c = 0; c < q; ++c) {
    long double t = 0.0L;
    for (int r = 0; r < s->n; ++r) {
      double v = s->X[(size_t)c * s->n + r];
      t += v * v;
    }
    xd[c] = (double)t;
  }
  double* pd = (double*)calloc(q, sizeof(double));
  unsigned char* pen = (unsigned char*)calloc(q, 1);
  double* sp = (double*)malloc(sizeof(double) * s->m);
  const double thr_scale = pow(DBL_EPSILON, 0.8);
  for (int i = 0; i < s->m; ++i) {
    int dim = s->dim[i], off = s->off[i];
    const double* B = s->S[i];
    double mx = 0.0;
    for (int e = 0; e < dim * dim; ++e) if (fabs(B[e]) > mx) mx = fabs(B[e]);
    double thr = thr_scale * mx;
    double* sx = (double*)malloc(sizeof(double) * dim);
    double* spn = (double*)malloc(sizeof(double) * dim);
    int cnt = 0;
    for (int r = 0; r < dim; ++r) {
      long double rt = 0.0L, ct = 0.0L;
      for (int o = 0; o < dim; ++o) {
        rt += fabs(B[(size_t)o * dim + r]);
        ct += fabs(B[(size_t)r * dim + o]);
      }
      double rmn = (double)(rt / (long double)dim), cmn = (double)(ct / (long double)dim);
      double dg = fabs(B[(size_t)r * dim + r]);
      if (rmn > thr && cmn > thr && dg > thr) {
        sx[cnt] = xd[off + r];
        spn[cnt] = B[(size_t)r * dim + r];
        pen[off + r] = 1;
        ++cnt;
      }
    }
    double mxd = orc_r_mean(sx, cnt, 1), msp = orc_r_mean(spn, cnt, 1);
    sp[i] = mxd / msp;
    for (int r = 0; r < dim; ++r) pd[off + r] += sp[i] * B[(size_t)r * dim + r];
    free(sx); free(spn);
  }
  double* vals = (double*)malloc(sizeof(double) * q);
  for (int guard = 0; guard < 4000; ++guard) {
    int cnt = 0;
    for (int c = 0; c < q; ++c)
      if (pd[c] > 0.0 && pen[c] && xd[c] > 0.0) vals[cnt++] = xd[c] / (xd[c] + pd[c]);
    double mh = orc_r_mean(vals, cnt, 1);
    if (mh > 0.4) {
      for (int i = 0; i < s->m; ++i) sp[i] *= 10.0;
      for (int c = 0; c < q; ++c) pd[c] *= 10.0;
    } else {
      break;
    }
  }
  for (int guard = 0; guard < 4000; ++guard) {
    int cnt = 0;
    for (int c = 0; c < q; ++c)
      if (pd[c] > 0.0 && pen[c] && xd[c] > 0.0) vals[cnt++] = xd[c] / (xd[c] + pd[c]);
    double mh = orc_r_mean(vals, cnt, 1);
    if (mh < 0.4) {
      for (int i = 0; i < s->m; ++i) sp[i] /= 10.0;
      for (int c = 0; c < q; ++c) pd[c] /= 10.0;
    } else {
      break;
    }
  }
  for (int i = 0; i < s->m; ++i) out[i] = log(sp[i]);
  free(xd); free(pd); free(pen); free(sp); free(vals);
}

static int magic_prepare(const orc_setup* s, double rank_tol, magic_problem* P) {
  memset(P, 0, sizeof(*P));
  int n = s->n, q = s->q, m = s->m;
  P->n = n; P->q = q; P->m = m; P->rank_tol = rank_tol; P->X = s->X;
  P->qr = (double*)malloc(sizeof(double) * (size_t)n * q);
  memcpy(P->qr, s->X, sizeof(double) * (size_t)n * q);
  P->tau = (double*)malloc(sizeof(double) * q);
  P->piv = (int*)malloc(sizeof(int) * q);
  orc_geqp3(P->qr, n, q, n, P->piv, P->tau);
  P->R = (double*)calloc((size_t)q * q, sizeof(double));
  for (int c = 0; c < q; ++c)
    for (int r = 0; r <= c; ++r) P->R[(size_t)c * q + r] = P->qr[(size_t)c * n + r];
  P->rS = (double**)calloc(m, sizeof(double*));
  P->Sp = (double**)calloc(m, sizeof(double*));
  P->rrank = (int*)malloc(sizeof(int) * m);
  for (int i = 0; i < m; ++i) {
    int dim = s->dim[i], rank = s->rank[i], off = s->off[i];
    double* br = block_root(s->S[i], dim, rank);
    if (!br) return -5;
    double* full = (double*)calloc((size_t)q * rank, sizeof(double));
    for (int r = 0; r < dim; ++r)
      for (int c = 0; c < rank; ++c) full[(size_t)c * q + off + r] = br[(size_t)c * dim + r];
    double* pr = (double*)malloc(sizeof(double) * (size_t)q * rank);
    for (int r = 0; r < q; ++r)
      for (int c = 0; c < rank; ++c) pr[(size_t)c * q + r] = full[(size_t)c * q + P->piv[r]];
    double* Sp = (double*)calloc((size_t)q * q, sizeof(double));
    for (int c = 0; c < q; ++c)
      for (int r = 0; r < q; ++r) {
        double v = 0.0;
        for (int k = 0; k < rank; ++k) v += pr[(size_t)k * q + r] * pr[(size_t)k * q + c];
        Sp[(size_t)c * q + r] = v;
      }
    P->rS[i] = pr;
    P->Sp[i] = Sp;
    P->rrank[i] = rank;
    free(br); free(full);
  }
  P->init_lsp = (double*)malloc(sizeof(double) * m);
  initial_lsp(s, P->init_lsp);
  return 0;
}

static void magic_problem_free(magic_problem* P) {
  free(P->R); free(P->qr); free(P->tau); free(P->piv);
  for (int i = 0; i < P->m; ++i) { free(P->rS[i]); free(P->Sp[i]); }
  free(P->rS); free(P->Sp); free(P->rrank); free(P->init_lsp);
}

static int magic_core_eval(const magic_problem* P, const double* y0, double ynorm2,
                           const double* lsp, magic_core* C) {
  int q = P->q;
  double* agg = (double*)calloc((size_t)q * q, sizeof(double));
  for (int i = 0; i < P->m; ++i) {
    double mult = exp(lsp[i]);
    for (int e = 0; e < q * q; ++e) agg[e] += P->Sp[i][e] * mult;
  }
  int* piv = (int*)malloc(sizeof(int) * q);
  int rank = 0;
  orc_pstf2_upper(agg, q, q, piv, -1.0, &rank);
  if (rank <= 0) { free(agg); free(piv); return -6; }
  int rows = q + rank;
  double* aug = (double*)calloc((size_t)rows * q, sizeof(double));
  for (int c = 0; c < q; ++c)
    for (int r = 0; r < q; ++r) aug[(size_t)c * rows + r] = P->R[(size_t)c * q + r];
  for (int col = 0; col < q; ++col) {
    int oc = piv[col];
    int cr = rank < col + 1 ? rank : col + 1;
    for (int r = 0; r < cr; ++r) aug[(size_t)oc * rows + q + r] = agg[(size_t)col * q + r];
  }
  double* U = (double*)malloc(sizeof(double) * (size_t)rows * q);
  double* d = (double*)malloc(sizeof(double) * q);
  double* V = (double*)malloc(sizeof(double) * (size_t)q * q);
  orc_svd_jacobi(aug, rows, q, U, d, V);
  double thr = d[0] * P->rank_tol;
  int nr = 0;
  C->rank_margin = 1e300;
  for (int i = 0; i < q; ++i) {
    if (d[i] >= thr) ++nr;
    double mg = fabs(d[i] / d[0] - P->rank_tol) / P->rank_tol;
    if (mg < C->rank_margin) C->rank_margin = mg;
  }
  C->arank = rank;
  C->nr = nr;
  C->U1 = (double*)malloc(sizeof(double) * (size_t)q * nr);
  C->V = (double*)malloc(sizeof(double) * (size_t)q * nr);
  C->d = (double*)malloc(sizeof(double) * nr);
  C->y1 = (double*)malloc(sizeof(double) * nr);
  for (int c = 0; c < nr; ++c) {
    for (int r = 0; r < q; ++r) {
      C->U1[(size_t)c * q + r] = U[(size_t)c * rows + r];
      C->V[(size_t)c * q + r] = V[(size_t)c * q + r];
    }
    C->d[c] = d[c];
  }
  double yAy = 0.0;
  for (int c = 0; c < nr; ++c) {
    double v = 0.0;
    for (int r = 0; r < q; ++r) v += C->U1[(size_t)c * q + r] * y0[r];
    C->y1[c] = v;
    yAy += v * v;
  }
  double yAAy = 0.0;
  for (int r = 0; r < q; ++r) {
    double v = 0.0;
    for (int c = 0; c < nr; ++c) v += C->U1[(size_t)c * q + r] * C->y1[c];
    yAAy += v * v;
  }
  C->rss = ynorm2 - 2.0 * yAy + yAAy;
  if (C->rss < 0.0) C->rss = 0.0;
  double edf = 0.0;
  for (int e = 0; e < q * nr; ++e) edf += C->U1[e] * C->U1[e];
  C->edf = edf;
  double delta = (double)P->n - edf;
  C->score = delta > 1e-8 ? (double)P->n * C->rss / (delta * delta) : INFINITY;
  C->cond = d[0] / d[nr - 1];
  free(agg); free(piv); free(aug); free(U); free(d); free(V);
  return 0;
}

/* Gradient and Hessian of the GCV score in log(sp). */
static void magic_derivs(const magic_problem* P, const magic_core* C, const double* lsp,
                         double* grad, double* hess) {
  int q = P->q, nr = C->nr, m = P->m;
  double* U1U1 = (double*)malloc(sizeof(double) * (size_t)nr * nr);
  for (int j = 0; j < nr; ++j)
    for (int i = 0; i < nr; ++i) {
      double v = 0.0;
      for (int r = 0; r < q; ++r) v += C->U1[(size_t)i * q + r] * C->U1[(size_t)j * q + r];
      U1U1[(size_t)j * nr + i] = v;
    }
  double** K = (double**)malloc(sizeof(double*) * m);
  double** Mm = (double**)malloc(sizeof(double*) * m);
  double** My = (double**)malloc(sizeof(double*) * m);
  double** Ky = (double**)malloc(sizeof(double*) * m);
  double** yK = (double**)malloc(sizeof(double*) * m);
  for (int i = 0; i < m; ++i) {
    int rk = P->rrank[i];
    const double* root = P->rS[i];
    double* tr = (double*)malloc(sizeof(double) * (size_t)nr * rk); /* nr x rk: V' root / d */
    for (int c = 0; c < rk; ++c)
      for (int r = 0; r < nr; ++r) {
        double v = 0.0;
        for (int k = 0; k < q; ++k) v += C->V[(size_t)r * q + k] * root[(size_t)c * q + k];
        tr[(size_t)c * nr + r] = v / C->d[r];
      }
    double* inter = (double*)malloc(sizeof(double) * (size_t)rk * nr); /* tr' U1U1 */
    for (int c = 0; c < nr; ++c)
      for (int r = 0; r < rk; ++r) {
        double v = 0.0;
        for (int k = 0; k < nr; ++k) v += tr[(size_t)r * nr + k] * U1U1[(size_t)c * nr + k];
        inter[(size_t)c * rk + r] = v;
      }
    K[i] = (double*)malloc(sizeof(double) * (size_t)nr * nr);
    Mm[i] = (double*)malloc(sizeof(double) * (size_t)nr * nr);
    for (int c = 0; c < nr; ++c)
      for (int r = 0; r < nr; ++r) {
        double v = 0.0, w = 0.0;
        for (int k = 0; k < rk; ++k) {
          v += tr[(size_t)k * nr + r] * inter[(size_t)c * rk + k];
          w += tr[(size_t)k * nr + r] * tr[(size_t)k * nr + c];
        }
        K[i][(size_t)c * nr + r] = v;
        Mm[i][(size_t)c * nr + r] = w;
      }
    My[i] = (double*)malloc(sizeof(double) * nr);
    Ky[i] = (double*)malloc(sizeof(double) * nr);
    yK[i] = (double*)malloc(sizeof(double) * nr);
    for (int o = 0; o < nr; ++o) {
      double mv = 0.0, yk = 0.0, ky = 0.0;
      for (int c = 0; c < nr; ++c) {
        mv += C->y1[c] * Mm[i][(size_t)o * nr + c];
        yk += C->y1[c] * K[i][(size_t)o * nr + c];
        ky += C->y1[c] * K[i][(size_t)c * nr + o];
      }
      My[i][o] = mv;
      yK[i][o] = yk;
      Ky[i][o] = ky;
    }
    free(tr); free(inter);
  }
  double* dnorm = (double*)malloc(sizeof(double) * m);
  double* ddelta = (double*)malloc(sizeof(double) * m);
  double* d2n = (double*)calloc((size_t)m * m, sizeof(double));
  double* d2d = (double*)calloc((size_t)m * m, sizeof(double));
  for (int i = 0; i < m; ++i) {
    double li = exp(lsp[i]);
    double trc = 0.0;
    for (int c = 0; c < nr; ++c) trc += K[i][(size_t)c * nr + c];
    ddelta[i] = li * trc;
    for (int j = 0; j <= i; ++j) {
      double ps = 0.0;
      for (int e = 0; e < nr * nr; ++e) ps += Mm[j][e] * K[i][e];
      double dv = -2.0 * exp(lsp[i] + lsp[j]) * ps;
      d2d[(size_t)j * m + i] = dv;
      d2d[(size_t)i * m + j] = dv;
    }
    d2d[(size_t)i * m + i] += ddelta[i];
    double nd = 0.0;
    for (int c = 0; c < nr; ++c) nd += C->y1[c] * (My[i][c] - Ky[i][c]);
    dnorm[i] = 2.0 * li * nd;
    for (int j = 0; j <= i; ++j) {
      double nv = 0.0;
      for (int c = 0; c < nr; ++c)
        nv += My[i][c] * Ky[j][c] + My[j][c] * Ky[i][c] - 2.0 * My[i][c] * My[j][c] + yK[i][c] * My[j][c];
      nv *= 2.0 * exp(lsp[i] + lsp[j]);
      d2n[(size_t)j * m + i] = nv;
      d2n[(size_t)i * m + j] = nv;
    }
    d2n[(size_t)i * m + i] += dnorm[i];
  }
  double delta = (double)P->n - C->edf;
  double ss = (double)P->n / (delta * delta);
  double ds = 2.0 * ss * C->rss / delta;
  double cs = -2.0 * ss / delta;
  double cv = 3.0 * ds / delta;
  for (int i = 0; i < m; ++i) grad[i] = ss * dnorm[i] - ds * ddelta[i];
  for (int i = 0; i < m; ++i)
    for (int j = 0; j <= i; ++j) {
      double v = cs * (ddelta[j] * dnorm[i] + ddelta[i] * dnorm[j]) + ss * d2n[(size_t)j * m + i] +
                 cv * ddelta[i] * ddelta[j] - ds * d2d[(size_t)j * m + i];
      hess[(size_t)j * m + i] = v;
      hess[(size_t)i * m + j] = v;
    }
  for (int i = 0; i < m; ++i) { free(K[i]); free(Mm[i]); free(My[i]); free(Ky[i]); free(yK[i]); }
  free(K); free(Mm); free(My); free(Ky); free(yK); free(U1U1);
  free(dnorm); free(ddelta); free(d2n); free(d2d);
}

static void magic_fit(const magic_problem* P, const magic_core* C, const double* y,
                      double* resid) {
  int q = P->q, n = P->n, nr = C->nr;
  double* theta = (double*)malloc(sizeof(double) * q);
  for (int r = 0; r < q; ++r) {
    double v = 0.0;
    for (int c = 0; c < nr; ++c) v += C->V[(size_t)c * q + r] * (C->y1[c] / C->d[c]);
    theta[r] = v;
  }
  double* beta = (double*)calloc(q, sizeof(double));
  for (int i = 0; i < q; ++i) beta[P->piv[i]] = theta[i];
  for (int r = 0; r < n; ++r) {
    double f = 0.0;
    for (int c = 0; c < q; ++c) f += P->X[(size_t)c * n + r] * beta[c];
    resid[r] = y[r] - f;
  }
  free(theta); free(beta);
}

int orc_magic_fit(const orc_setup* s, const double* y, const orc_magic_control* ctl,
                  double* resid, orc_magic_result* res) {
  magic_problem P;
  int rc = magic_prepare(s, ctl->rank_tol, &P);
  if (rc != 0) { magic_problem_free(&P); return rc; }
  int n = s->n, q = s->q, m = s->m;
  double* y0 = (double*)malloc(sizeof(double) * n);
  memcpy(y0, y, sizeof(double) * n);
  orc_apply_qt(P.qr, n, q, n, P.tau, y0);
  double yn = 0.0;
  for (int i = 0; i < n; ++i) yn += y[i] * y[i];
  double* lsp = (double*)malloc(sizeof(double) * m);
  double* trial = (double*)malloc(sizeof(double) * m);
  double* grad = (double*)calloc(m, sizeof(double));
  double* hess = (double*)calloc((size_t)m * m, sizeof(double));
  double* cgrad = (double*)calloc(m, sizeof(double));
  double* nstep = (double*)calloc(m, sizeof(double));
  double* sstep = (double*)calloc(m, sizeof(double));
  double* step = (double*)calloc(m, sizeof(double));
  double* ev = (double*)malloc(sizeof(double) * m);
  double* evec = (double*)malloc(sizeof(double) * (size_t)m * m);
  memcpy(lsp, P.init_lsp, sizeof(double) * m);
  const int max_half = m == 1 ? ctl->max_step_halving_single : ctl->max_step_halving_multi;
  magic_core cur;
  memset(&cur, 0, sizeof(cur));
  rc = magic_core_eval(&P, y0, yn, lsp, &cur);
  if (rc != 0) goto done;
  magic_derivs(&P, &cur, lsp, cgrad, hess);
  double min_score = cur.score, score_red = 1e10;
  /* smallest normalised distance of any optimizer decision from its
   * threshold; below ~1e-10 a rounding difference can flip the decision */
  double margin = cur.rank_margin;
#define ORC_MARGIN(v) do { double mv_ = (v); if (mv_ < margin) margin = mv_; } while (0)
  int use_sd = 0, converged = 0, iter, step_failed = 0, evals = 1;
  for (iter = 1; iter <= ctl->max_iterations; ++iter) {
    int tries = 0;
    if (iter > 1) {
      memcpy(step, use_sd ? sstep : nstep, sizeof(double) * m);
      int trying = 1;
      while (trying) {
        ++tries;
        if (tries == 4 && !use_sd) { use_sd = 1; memcpy(step, sstep, sizeof(double) * m); }
        for (int i = 0; i < m; ++i) trial[i] = lsp[i] + step[i];
        magic_core tc;
        memset(&tc, 0, sizeof(tc));
        rc = magic_core_eval(&P, y0, yn, trial, &tc);
        ++evals;
        if (rc != 0) goto done;
        ORC_MARGIN(tc.rank_margin);
        ORC_MARGIN(fabs(tc.score - min_score) / min_score);
        if (tc.score < min_score) {
          trying = 0;
          score_red = min_score - tc.score;
          min_score = tc.score;
          memcpy(lsp, trial, sizeof(double) * m);
          core_free(&cur);
          cur = tc;
          magic_derivs(&P, &cur, lsp, cgrad, hess);
        } else {
          core_free(&tc);
          for (int i = 0; i < m; ++i) step[i] /= 2.0;
        }
        if (tries == max_half - 1 && trying) for (int i = 0; i < m; ++i) step[i] = 0.0;
        if (tries == max_half) trying = 0;
      }
    }
    if (iter > 3) {
      double gn = 0.0;
      for (int i = 0; i < m; ++i) gn += grad[i] * grad[i];
      gn = sqrt(gn);
      converged = score_red <= ctl->tol * (1.0 + min_score) &&
                  gn <= pow(ctl->tol, 1.0 / 3.0) * (1.0 + fabs(min_score));
      {
        double t1 = ctl->tol * (1.0 + min_score);
        double t2 = pow(ctl->tol, 1.0 / 3.0) * (1.0 + fabs(min_score));
        ORC_MARGIN(fabs(score_red - t1) / t1);
        ORC_MARGIN(fabs(gn - t2) / t2);
      }
      if (tries == max_half) { converged = 1; step_failed = 1; }
    }
    memcpy(grad, cgrad, sizeof(double) * m);
    orc_syev_jacobi(hess, m, ev, evec);
    use_sd = 0;
    for (int i = 0; i < m; ++i) {
      if (ev[i] < 0.0) use_sd = 1;
      ORC_MARGIN(fabs(ev[i]) / min_score);
    }
    if (!use_sd) {
      for (int r = 0; r < m; ++r) {
        double v = 0.0;
        for (int c = 0; c < m; ++c) {
          double pc = 0.0;
          for (int k = 0; k < m; ++k) pc += evec[(size_t)c * m + k] * grad[k];
          v += evec[(size_t)c * m + r] * (pc / ev[c]);
        }
        nstep[r] = -v;
      }
      double mx = fabs(nstep[0]);
      for (int i = 1; i < m; ++i) if (fabs(nstep[i]) > mx) mx = fabs(nstep[i]);
      if (mx > ctl->max_newton_step)
        for (int i = 0; i < m; ++i) nstep[i] *= ctl->max_newton_step / mx;
    }
    double mg = fabs(grad[0]);
    for (int i = 1; i < m; ++i) if (fabs(grad[i]) > mg) mg = fabs(grad[i]);
    for (int i = 0; i < m; ++i) sstep[i] = mg == 0.0 ? 0.0 : -grad[i] / mg;
    if (converged) break;
  }
  if (!converged) { rc = -7; goto done; }
  res->iterations = iter;
  res->step_failed = step_failed;
  for (int c = 0; c < m; ++c) {
    double dir = grad[c] < 0.0 ? 1.0 : -1.0;
    for (int pr = 1; pr <= ctl->max_boundary_probes; ++pr) {
      memcpy(trial, lsp, sizeof(double) * m);
      trial[c] += dir * ctl->boundary_probe_step;
      magic_core tc;
      memset(&tc, 0, sizeof(tc));
      rc = magic_core_eval(&P, y0, yn, trial, &tc);
      ++evals;
      if (rc != 0) goto done;
      ORC_MARGIN(tc.rank_margin);
      ORC_MARGIN(fabs(tc.score - min_score) / min_score);
      if (!(tc.score < min_score)) { core_free(&tc); break; }
      memcpy(lsp, trial, sizeof(double) * m);
      core_free(&cur);
      cur = tc;
      min_score = cur.score;
    }
  }
  /* final fit at the selected log(sp) */
  {
    magic_core fin;
    memset(&fin, 0, sizeof(fin));
    rc = magic_core_eval(&P, y0, yn, lsp, &fin);
    if (rc != 0) goto done;
    magic_fit(&P, &fin, y, resid);
    res->score = fin.score;
    res->rss = fin.rss;
    res->edf = fin.edf;
    core_free(&fin);
  }
  for (int i = 0; i < m && i < 64; ++i) res->log_sp[i] = lsp[i];
  res->evaluations = evals;
  res->margin = margin;
  rc = 0;
done:
  core_free(&cur);
  free(y0); free(lsp); free(trial); free(grad); free(hess); free(cgrad); free(nstep);
  free(sstep); free(step); free(ev); free(evec);
  magic_problem_free(&P);
  return rc;
}

/* GCV score, gradient and Hessian at a given log(sp) (used to pin the
 * derivative formulas against finite differences). */
int orc_magic_eval(const orc_setup* s, const double* y, const double* lsp, double rank_tol,
                   double* score, double* grad, double* hess) {
  magic_problem P;
  int rc = magic_prepare(s, rank_tol, &P);
  if (rc != 0) { magic_problem_free(&P); return rc; }
  double* y0 = (double*)malloc(sizeof(double) * s->n);
  memcpy(y0, y, sizeof(double) * s->n);
  orc_apply_qt(P.qr, s->n, s->q, s->n, P.tau, y0);
  double yn = 0.0;
  for (int i = 0; i < s->n; ++i) yn += y[i] * y[i];
  magic_core c;
  memset(&c, 0, sizeof(c));
  rc = magic_core_eval(&P, y0, yn, lsp, &c);
  if (rc == 0) {
    *score = c.score;
    if (grad && hess) magic_derivs(&P, &c, lsp, grad, hess);
  }
  core_free(&c);
  free(y0);
  magic_problem_free(&P);
  return rc;
}

void orc_magic_default_control(orc_magic_control* c) {
  c->tol = 1e-7;
  c->max_step_halving_single = 15;
  c->max_step_halving_multi = 25;
  c->max_iterations = 400;
  c->max_newton_step = 5.0;
  c->boundary_probe_step = 2.0;
  c->max_boundary_probes = 5;
  c->rank_tol = 1.4901161193847656e-8;
}

int orc_regrxons(const double* data, int n, int ld, int target, const int* S, int s_len,
                 const orc_magic_control* ctl, double* resid, orc_magic_result* res) {
  orc_setup s;
  int rc = orc_regr_setup(data, n, ld, S, s_len, &s);
  if (rc != 0) return rc;
  rc = orc_magic_fit(&s, data + (size_t)target * ld, ctl, resid, res);
  orc_setup_free(&s);
  return rc;
}

int orc_regrxons_at_sp(const double* data, int n, int ld, int target, const int* S, int s_len,
                       const double* log_sp, double rank_tol, double* resid, double out[3]) {
  orc_setup s;
  int rc = orc_regr_setup(data, n, ld, S, s_len, &s);
  if (rc != 0) return rc;
  magic_problem P;
  rc = magic_prepare(&s, rank_tol, &P);
  if (rc == 0) {
    const double* y = data + (size_t)target * ld;
    double* y0 = (double*)malloc(sizeof(double) * n);
    memcpy(y0, y, sizeof(double) * n);
    orc_apply_qt(P.qr, n, s.q, n, P.tau, y0);
    double yn = 0.0;
    for (int i = 0; i < n; ++i) yn += y[i] * y[i];
    magic_core c;
    memset(&c, 0, sizeof(c));
    rc = magic_core_eval(&P, y0, yn, log_sp, &c);
    if (rc == 0) {
      magic_fit(&P, &c, y, resid);
      out[0] = c.score;
      out[1] = c.rss;
      out[2] = c.edf;
    }
    core_free(&c);
    free(y0);
  }
  magic_problem_free(&P);
  orc_setup_free(&s);
  return rc;
}
