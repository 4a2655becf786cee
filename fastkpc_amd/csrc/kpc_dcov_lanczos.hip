// dcov.gamma components by Lanczos on the implicit distance matrix.
//
// Same output contract as dcov_component_kernel (top-k |lambda| eigenpairs
// of D_ij = |r_i - r_j| selected with the higher-index tie break, centred
// vectors, sum(D) and the self moment; kpcalg/R/dcovgamma.R:67-79,
// fastkpc/src/legacy_dcov_gamma_cpp.cpp:60-78, :135, :283-300), but D is
// never stored: every Lanczos step recomputes D v from the residual held in
// LDS (n^2 FMAs, no memory traffic).  Full reorthogonalisation (two
// classical Gram-Schmidt passes) keeps the Krylov basis orthogonal, and the
// iteration stops when every selected Ritz pair has residual
// |beta_m s_m| <= tol * ||T||; the extreme eigenvalues of the Lanczos
// tridiagonal T come from Sturm bisection and their eigenvectors from
// inverse iteration (one lane per eigenpair).  Kpcalg's own eigs() call
// (RSpectra, tol 1e-10) is a restarted Lanczos of the same operator.
#include "kpc_common.h"
#include "kpc_device.h"

namespace kpc {

constexpr int kLzThreads = 512;

__device__ int lz_sturm(const double* d, const double* e2, int n, double x, double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  cnt += q < 0.0;
  for (int i = 1; i < n; ++i) {
    q = d[i] - x - e2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
  }
  return cnt;
}

__device__ double lz_bisect(const double* d, const double* e2, int n, int idx, double lo, double hi,
                            double pivmin) {
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (hi - lo <= 2.0 * 2.220446049250313e-16 * fmax(fabs(lo), fabs(hi)) + pivmin) break;
    if (lz_sturm(d, e2, n, mid, pivmin) > idx) hi = mid; else lo = mid;
  }
  return 0.5 * (lo + hi);
}

// Inverse iteration on the m x m tridiagonal (d, e) for eigenvalue lam;
// ws: 5m doubles; z (m) receives the unit eigenvector.
__device__ void lz_invit(const double* d, const double* e, int m, double lam, double tnorm,
                         double* ws, double* z) {
  double* dl = ws;
  double* dm = ws + m;
  double* du = ws + 2 * m;
  double* du2 = ws + 3 * m;
  int* ip = reinterpret_cast<int*>(ws + 4 * m);
  const double pert = 2.220446049250313e-16 * tnorm;
  for (int i = 0; i < m; ++i) dm[i] = d[i] - lam;
  for (int i = 0; i < m - 1; ++i) { dl[i] = e[i]; du[i] = e[i]; du2[i] = 0.0; }
  for (int i = 0; i < m - 1; ++i) {
    if (fabs(dm[i]) >= fabs(dl[i])) {
      ip[i] = 0;
      if (dm[i] == 0.0) dm[i] = pert;
      const double f = dl[i] / dm[i];
      dl[i] = f;
      dm[i + 1] -= f * du[i];
    } else {
      ip[i] = 1;
      const double f = dm[i] / dl[i];
      dm[i] = dl[i];
      dl[i] = f;
      const double t = du[i];
      du[i] = dm[i + 1];
      dm[i + 1] = t - f * dm[i + 1];
      if (i < m - 2) { du2[i] = du[i + 1]; du[i + 1] = -f * du[i + 1]; }
    }
  }
  if (dm[m - 1] == 0.0) dm[m - 1] = pert;
  for (int i = 0; i < m; ++i) z[i] = 1.0 + 0.01 * (double)((i * 7919) % 101) / 101.0;
  for (int it = 0; it < 3; ++it) {
    for (int i = 0; i < m - 1; ++i) {
      if (ip[i] == 0) {
        z[i + 1] -= dl[i] * z[i];
      } else {
        const double t = z[i];
        z[i] = z[i + 1];
        z[i + 1] = t - dl[i] * z[i];
      }
    }
    z[m - 1] /= dm[m - 1];
    if (m >= 2) z[m - 2] = (z[m - 2] - du[m - 2] * z[m - 1]) / dm[m - 2];
    for (int i = m - 3; i >= 0; --i) z[i] = (z[i] - du[i] * z[i + 1] - du2[i] * z[i + 2]) / dm[i];
    double nrm = 0.0;
    for (int i = 0; i < m; ++i) nrm += z[i] * z[i];
    nrm = 1.0 / sqrt(nrm);
    for (int i = 0; i < m; ++i) z[i] *= nrm;
  }
}

__global__ __launch_bounds__(kLzThreads) void dcov_lanczos_kernel(
    const double* __restrict__ src_base, const long long* __restrict__ src_off,
    const int* __restrict__ dst_slot, int count, int n, int k, double* __restrict__ comp_vals,
    double* __restrict__ comp_vecs, double* __restrict__ comp_dsum,
    double* __restrict__ comp_moment, int* __restrict__ comp_status, double* __restrict__ scratch,
    long long scratch_stride, int mmax, double tol) {
  extern __shared__ double lds_dyn[];
  const int nl = (n + 63) & ~63;
  const int ml = (mmax + 64) & ~63;
  double* r = lds_dyn;                 // n
  double* qv = r + nl;                 // current Lanczos vector
  double* z = qv + nl;                 // residual vector
  double* al = z + nl;                 // ml
  double* be = al + ml;                // ml
  double* e2 = be + ml;                // ml
  double* cg = e2 + ml;                // ml (CGS coefficients)
  double* red = cg + ml;               // 16
  double* cval = red + 16;             // 2 kMaxNumCol
  double* sval = cval + 2 * kMaxNumCol;   // kMaxNumCol
  double* sres = sval + kMaxNumCol;       // kMaxNumCol
  int* cidx = reinterpret_cast<int*>(sres + kMaxNumCol);  // 2 kMaxNumCol
  int* sidx = cidx + 2 * kMaxNumCol;                      // kMaxNumCol
  int* ictl = sidx + kMaxNumCol;                          // 8
  const int tid = threadIdx.x;
  const int lane = tid & (KPC_WAVE - 1);
  const int wid = tid / KPC_WAVE;
  const int nw = blockDim.x / KPC_WAVE;
  double* Q = scratch + (size_t)blockIdx.x * scratch_stride;   // n x (mmax + 1)
  double* S = Q + (size_t)n * (mmax + 1);                      // k x mmax (Ritz coefficients)
  double* ivws = S + (size_t)k * mmax;                         // k x 5 mmax
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const double* x = src_base + src_off[item];
    const int slot = dst_slot[item];
    for (int i = tid; i < n; i += blockDim.x) r[i] = x[i];
    __syncthreads();
    // sum(D) in double-double
    dd acc{0.0, 0.0};
    for (int i = tid; i < n; i += blockDim.x) {
      const double ri = r[i];
      dd rowacc{0.0, 0.0};
      for (int j = 0; j < n; ++j) rowacc = dd_add(rowacc, dd{fabs(ri - r[j]), 0.0});
      acc = dd_add(acc, rowacc);
    }
    acc = block_sum_dd(acc, red);
    const double dsum = acc.hi + acc.lo;
    // start vector: mgcv-style LCG values minus 0.5, normalised
    if (tid == 0) {
      unsigned long st = 1;
      double nrm = 0.0;
      for (int i = 0; i < n; ++i) {
        st = (st * 106UL + 1283UL) % 6075UL;
        const double v = (double)st / 6075.0 - 0.5;
        qv[i] = v;
        nrm += v * v;
      }
      red[8] = 1.0 / sqrt(nrm);
    }
    __syncthreads();
    const double inv0 = red[8];
    __syncthreads();
    for (int i = tid; i < n; i += blockDim.x) {
      qv[i] *= inv0;
      Q[i] = qv[i];
    }
    __syncthreads();
    int m_final = mmax;
    if (tid == 0) { ictl[1] = 0; ictl[2] = 2; }
    __syncthreads();
    for (int j = 0; j < mmax; ++j) {
      // z = D q_j, recomputed from r
      double part = 0.0;
      for (int i = tid; i < n; i += blockDim.x) {
        const double ri = r[i];
        double s = 0.0;
        for (int l = 0; l < n; ++l) s += fabs(ri - r[l]) * qv[l];
        z[i] = s;
        part += qv[i] * s;
      }
      const double aj = block_sum(part, red);
      if (tid == 0) al[j] = aj;
      const double bprev = j > 0 ? be[j - 1] : 0.0;
      const double* qp = j > 0 ? Q + (size_t)(j - 1) * n : Q;
      for (int i = tid; i < n; i += blockDim.x) z[i] -= aj * qv[i] + bprev * qp[i];
      __syncthreads();
      for (int pass = 0; pass < 2; ++pass) {
        for (int b = wid; b <= j; b += nw) {
          const double* qb = Q + (size_t)b * n;
          double s = 0.0;
          for (int i = lane; i < n; i += KPC_WAVE) s += z[i] * qb[i];
          s = wave_sum(s);
          if (lane == 0) cg[b] = s;
        }
        __syncthreads();
        for (int i = tid; i < n; i += blockDim.x) {
          double s = 0.0;
          for (int b = 0; b <= j; ++b) s += cg[b] * Q[(size_t)b * n + i];
          z[i] -= s;
        }
        __syncthreads();
      }
      part = 0.0;
      for (int i = tid; i < n; i += blockDim.x) part += z[i] * z[i];
      const double bj = sqrt(block_sum(part, red));
      if (tid == 0) be[j] = bj;
      const double inv = bj > 1e-300 ? 1.0 / bj : 0.0;
      for (int i = tid; i < n; i += blockDim.x) {
        const double v = z[i] * inv;
        qv[i] = v;
        Q[(size_t)(j + 1) * n + i] = v;
      }
      __syncthreads();
      const int m = j + 1;
      const bool check = (m >= k + 1 && ((m - (k + 1)) % 5 == 0)) || m == mmax || bj <= 1e-300;
      if (!check) continue;
      // extreme eigenvalues of T (ascending indices {0..k-1} U {max(k, m-k)..m-1})
      for (int i = tid; i < m - 1; i += blockDim.x) e2[i] = be[i] * be[i];
      double gl = 1e308, gu = -1e308, tn = 0.0, emx = 0.0;
      for (int i = tid; i < m; i += blockDim.x) {
        const double rad = (i > 0 ? fabs(be[i - 1]) : 0.0) + (i < m - 1 ? fabs(be[i]) : 0.0);
        gl = fmin(gl, al[i] - rad);
        gu = fmax(gu, al[i] + rad);
        tn = fmax(tn, fabs(al[i]) + rad);
        if (i < m - 1) emx = fmax(emx, be[i] * be[i]);
      }
      gl = -block_max(-gl, red);
      gu = block_max(gu, red);
      tn = block_max(tn, red);
      emx = block_max(emx, red);
      const double pivmin = 2.2250738585072014e-308 * fmax(1.0, emx);
      const double span = gu - gl;
      const double lo0 = gl - 2.0 * 2.220446049250313e-16 * span - pivmin;
      const double hi0 = gu + 2.0 * 2.220446049250313e-16 * span + pivmin;
      const int kk = k < m ? k : m;
      const int nlow = kk;
      const int hs = (m - kk) > nlow ? (m - kk) : nlow;
      const int ncand = nlow + (m - hs);
      for (int c = tid; c < ncand; c += blockDim.x) {
        const int idx = c < nlow ? c : hs + (c - nlow);
        cidx[c] = idx;
        cval[c] = lz_bisect(al, e2, m, idx, lo0, hi0, pivmin);
      }
      __syncthreads();
      if (tid == 0) {
        for (int s = 0; s < kk; ++s) {
          int best = -1;
          for (int c = 0; c < ncand; ++c) {
            if (cidx[c] < 0) continue;
            if (best < 0) { best = c; continue; }
            const double a = fabs(cval[c]), b = fabs(cval[best]);
            if (a > b || (a == b && cidx[c] > cidx[best])) best = c;
          }
          sval[s] = cval[best];
          sidx[s] = cidx[best];
          cidx[best] = -1;
        }
      }
      __syncthreads();
      for (int c = tid; c < kk; c += blockDim.x) {
        double* sv = S + (size_t)c * mmax;
        lz_invit(al, be, m, sval[c], tn, ivws + (size_t)c * 5 * mmax, sv);
        sres[c] = fabs(bj * sv[m - 1]);
      }
      __syncthreads();
      if (tid == 0) {
        double worst = 0.0;
        for (int c = 0; c < kk; ++c) worst = fmax(worst, sres[c]);
        const double tnorm = fmax(fabs(lo0), fabs(hi0));
        const int conv = kk == k && worst <= tol * tnorm;
        ictl[0] = conv || m == mmax || bj <= 1e-300;
        ictl[1] = kk;
        // status: 0 converged, 1 Krylov space exhausted (exact), 2 step cap hit
        ictl[2] = conv ? 0 : (bj <= 1e-300 ? 1 : 2);
      }
      __syncthreads();
      if (ictl[0]) { m_final = m; break; }
    }
    // fewer than k eigenpairs when the Krylov space was exhausted
    for (int c = ictl[1] + tid; c < k; c += blockDim.x) {
      sval[c] = 0.0;
      double* sv = S + (size_t)c * mmax;
      for (int b = 0; b < m_final; ++b) sv[b] = 0.0;
    }
    __syncthreads();
    // Ritz vectors U = Q S -> output
    double* Z = comp_vecs + (size_t)slot * n * k;
    for (int e = tid; e < n * k; e += blockDim.x) {
      const int i = e % n, c = e / n;
      const double* sv = S + (size_t)c * mmax;
      double s = 0.0;
      for (int b = 0; b < m_final; ++b) s += Q[(size_t)b * n + i] * sv[b];
      Z[(size_t)c * n + i] = s;
    }
    __syncthreads();
    for (int c = wid; c < k; c += nw) {
      double* zc = Z + (size_t)c * n;
      double s = 0.0;
      for (int i = lane; i < n; i += KPC_WAVE) s += zc[i];
      s = wave_sum(s) / (double)n;
      for (int i = lane; i < n; i += KPC_WAVE) zc[i] -= s;
    }
    __syncthreads();
    double mpart = 0.0;
    for (int e = tid; e < k * k; e += blockDim.x) {
      const int i = e % k, jj = e / k;
      const double* ci = Z + (size_t)i * n;
      const double* cj = Z + (size_t)jj * n;
      double s = 0.0;
      for (int rr = 0; rr < n; ++rr) s += ci[rr] * cj[rr];
      mpart += sval[i] * sval[jj] * s * s;
    }
    const double moment = block_sum(mpart, red);
    if (tid == 0) {
      comp_dsum[slot] = dsum;
      comp_moment[slot] = moment;
      // Lanczos steps taken; negative when the step cap stopped an
      // unconverged iteration short of n (inexact)
      comp_status[slot] = (ictl[2] == 2 && m_final < n) ? -m_final : m_final;
    }
    for (int c = tid; c < k; c += blockDim.x) comp_vals[(size_t)slot * k + c] = sval[c];
    __syncthreads();
  }
}

int lanczos_mmax(int n) { return n < 1024 ? n : 1024; }

size_t dcov_lanczos_lds_bytes(int n, int mmax) {
  const int nl = (n + 63) & ~63;
  const int ml = (mmax + 64) & ~63;
  return (size_t)(3 * nl + 4 * ml + 16 + 4 * kMaxNumCol) * sizeof(double) +
         (size_t)(3 * kMaxNumCol + 8) * sizeof(int);
}

long long dcov_lanczos_scratch_doubles(int n, int k, int mmax) {
  long long s = (long long)n * (mmax + 1) + (long long)k * mmax + (long long)k * 5 * mmax + 64;
  return (s + 31) / 32 * 32;
}

hipError_t launch_dcov_lanczos(const double* src_base, const long long* src_off,
                               const int* dst_slot, int count, int n, int k, double* vals,
                               double* vecs, double* dsum, double* moment, int* status,
                               double* scratch, int grid, int mmax, double tol,
                               hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  const size_t lds = dcov_lanczos_lds_bytes(n, mmax);
  hipError_t err = hipFuncSetAttribute((const void*)dcov_lanczos_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(dcov_lanczos_kernel, dim3(grid), dim3(kLzThreads), lds, stream, src_base,
                     src_off, dst_slot, count, n, k, vals, vecs, dsum, moment, status, scratch,
                     dcov_lanczos_scratch_doubles(n, k, mmax), mmax, tol);
  return hipGetLastError();
}

}  // namespace kpc
