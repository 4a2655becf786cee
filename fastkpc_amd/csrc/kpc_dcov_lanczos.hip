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
// tridiagonal T come from Sturm-count multisection (a group of lanes per
// eigenvalue, each probing one point of the bracket per round) and their
// eigenvectors from twisted factorisations (one lane per eigenpair, pivots in
// LDS).  The convergence checks are scheduled from the observed geometric
// decay of the worst Ritz residual; a beta below 64 eps |T| is a breakdown
// (invariant Krylov space, exact Ritz pairs) and ends the iteration.  Kpcalg's own eigs() call (RSpectra, tol 1e-10) is a
// restarted Lanczos of the same operator.
//
// Dependent FP64 operations cost tens of cycles each on CDNA4, so every
// serial recurrence here uses hardware reciprocal estimates refined by
// Newton steps and every long dot product runs several accumulators.
#include "kpc_common.h"
#include "kpc_device.h"
#include "kpc_tridiag.h"

namespace kpc {

constexpr int kLzThreads = 512;

// Doubles of LDS for the twisted-factorisation pivots (2 m per eigenpair).
__host__ __device__ inline int lanczos_ivw(int mmax) {
  return 4096 > 6 * mmax + 64 ? 4096 : 6 * mmax + 64;
}

// Top-kk candidates by |value| (ties toward the higher ascending index),
// lane 0: sval/sidx receive values and candidate slots.
__device__ void lz_select(const double* cval, int ncand, int nlow, int hs, int kk, double* sval,
                          int* sidx, int* used) {
  if (threadIdx.x == 0) {
    for (int c = 0; c < ncand; ++c) used[c] = 0;
    for (int s2 = 0; s2 < kk; ++s2) {
      int best = -1;
      for (int c = 0; c < ncand; ++c) {
        if (used[c]) continue;
        if (best < 0) { best = c; continue; }
        const double a = fabs(cval[c]), b = fabs(cval[best]);
        const int ic = c < nlow ? c : hs + (c - nlow), ib = best < nlow ? best : hs + (best - nlow);
        if (a > b || (a == b && ic > ib)) best = c;
      }
      sval[s2] = cval[best];
      sidx[s2] = best;
      used[best] = 1;
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(kLzThreads) void dcov_lanczos_kernel(
    const double* __restrict__ src_base, const long long* __restrict__ src_off,
    const int* __restrict__ dst_slot, int count, int n, int k, double* __restrict__ comp_vals,
    double* __restrict__ comp_vecs, double* __restrict__ comp_dsum,
    double* __restrict__ comp_moment, int* __restrict__ comp_status, double* __restrict__ scratch,
    long long scratch_stride, int mmax, double tol) {
  extern __shared__ double2 lds_v2[];
  double* lds_dyn = reinterpret_cast<double*>(lds_v2);
  const int nl = (n + 63) & ~63;
  const int ml = (mmax + 64) & ~63;
  double2* rq = lds_v2;                // n: (residual, current Lanczos vector)
  double* z = lds_dyn + 2 * nl;        // residual vector
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
  double* clo = reinterpret_cast<double*>(ictl + 8);      // 2 kMaxNumCol brackets
  double* chi = clo + 2 * kMaxNumCol;                     // 2 kMaxNumCol
  double* ivl = chi + 2 * kMaxNumCol;                     // inverse-iteration workspace
  const int ivw = lanczos_ivw(mmax);
  const int tid = threadIdx.x;
  const int lane = tid & (KPC_WAVE - 1);
  const int wid = tid / KPC_WAVE;
  const int nw = blockDim.x / KPC_WAVE;
  double* Q = scratch + (size_t)blockIdx.x * scratch_stride;   // n x (mmax + 1)
  double* S = Q + (size_t)n * (mmax + 1);                      // k x mmax (Ritz coefficients)
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const double* x = src_base + src_off[item];
    const int slot = dst_slot[item];
    for (int i = tid; i < n; i += blockDim.x) rq[i].x = x[i];
    __syncthreads();
    // sum(D) in double-double
    dd acc{0.0, 0.0};
    for (int i = tid; i < n; i += blockDim.x) {
      const double ri = rq[i].x;
      dd rowacc{0.0, 0.0};
      for (int j = 0; j < n; ++j) rowacc = dd_add(rowacc, dd{fabs(ri - rq[j].x), 0.0});
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
        z[i] = v;
        nrm += v * v;
      }
      red[8] = 1.0 / sqrt(nrm);
    }
    __syncthreads();
    const double inv0 = red[8];
    __syncthreads();
    for (int i = tid; i < n; i += blockDim.x) {
      const double v = z[i] * inv0;
      rq[i].y = v;
      Q[i] = v;
    }
    __syncthreads();
    int m_final = mmax;
    int next_check = k + 1, prev_m = 0;
    double prev_worst = 0.0;
    double tmax = 0.0;  // running max of |alpha_j| + |beta_j| (scale of T)
    if (tid == 0) { ictl[1] = 0; ictl[2] = 2; }
    __syncthreads();
    for (int j = 0; j < mmax; ++j) {
      // z = D q_j recomputed from the residual, eight accumulators per row
      double part = 0.0;
      for (int i = tid; i < n; i += blockDim.x) {
        const double ri = rq[i].x;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, a4 = 0.0, a5 = 0.0, a6 = 0.0, a7 = 0.0;
        int l = 0;
        for (; l + 8 <= n; l += 8) {
          const double2 v0 = rq[l], v1 = rq[l + 1], v2 = rq[l + 2], v3 = rq[l + 3];
          const double2 v4 = rq[l + 4], v5 = rq[l + 5], v6 = rq[l + 6], v7 = rq[l + 7];
          a0 += fabs(ri - v0.x) * v0.y;
          a1 += fabs(ri - v1.x) * v1.y;
          a2 += fabs(ri - v2.x) * v2.y;
          a3 += fabs(ri - v3.x) * v3.y;
          a4 += fabs(ri - v4.x) * v4.y;
          a5 += fabs(ri - v5.x) * v5.y;
          a6 += fabs(ri - v6.x) * v6.y;
          a7 += fabs(ri - v7.x) * v7.y;
        }
        for (; l < n; ++l) a0 += fabs(ri - rq[l].x) * rq[l].y;
        const double sacc = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
        z[i] = sacc;
        part += rq[i].y * sacc;
      }
      const double aj = block_sum(part, red);
      if (tid == 0) al[j] = aj;
      const double bprev = j > 0 ? be[j - 1] : 0.0;
      const double* qp = j > 0 ? Q + (size_t)(j - 1) * n : Q;
      part = 0.0;
      for (int i = tid; i < n; i += blockDim.x) {
        const double v = z[i] - (aj * rq[i].y + bprev * qp[i]);
        z[i] = v;
        part += v * v;
      }
      double zn2 = block_sum(part, red);
      // classical Gram-Schmidt against Q[0..j], repeated while the norm drops
      // below 1/sqrt(2) of its value before the pass (DGKS, at most twice); a
      // wave forms eight coefficients at once (lanes over rows, coalesced),
      // then every row subtracts its combination with four accumulators
      for (int pass = 0; pass < 2; ++pass) {
        for (int b0 = wid * 8; b0 <= j; b0 += nw * 8) {
          double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0, c4 = 0.0, c5 = 0.0, c6 = 0.0, c7 = 0.0;
          const int nb = min(8, j + 1 - b0);
          const double* qb = Q + (size_t)b0 * n;
          for (int i = lane; i < n; i += KPC_WAVE) {
            const double zi = z[i];
            c0 += zi * qb[i];
            if (nb > 1) c1 += zi * qb[(size_t)n + i];
            if (nb > 2) c2 += zi * qb[(size_t)2 * n + i];
            if (nb > 3) c3 += zi * qb[(size_t)3 * n + i];
            if (nb > 4) c4 += zi * qb[(size_t)4 * n + i];
            if (nb > 5) c5 += zi * qb[(size_t)5 * n + i];
            if (nb > 6) c6 += zi * qb[(size_t)6 * n + i];
            if (nb > 7) c7 += zi * qb[(size_t)7 * n + i];
          }
          c0 = wave_sum(c0);
          c1 = wave_sum(c1);
          c2 = wave_sum(c2);
          c3 = wave_sum(c3);
          c4 = wave_sum(c4);
          c5 = wave_sum(c5);
          c6 = wave_sum(c6);
          c7 = wave_sum(c7);
          if (lane == 0) {
            cg[b0] = c0;
            if (nb > 1) cg[b0 + 1] = c1;
            if (nb > 2) cg[b0 + 2] = c2;
            if (nb > 3) cg[b0 + 3] = c3;
            if (nb > 4) cg[b0 + 4] = c4;
            if (nb > 5) cg[b0 + 5] = c5;
            if (nb > 6) cg[b0 + 6] = c6;
            if (nb > 7) cg[b0 + 7] = c7;
          }
        }
        __syncthreads();
        for (int i = tid; i < n; i += blockDim.x) {
          double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
          int b = 0;
          for (; b + 4 <= j + 1; b += 4) {
            s0 += cg[b] * Q[(size_t)b * n + i];
            s1 += cg[b + 1] * Q[(size_t)(b + 1) * n + i];
            s2 += cg[b + 2] * Q[(size_t)(b + 2) * n + i];
            s3 += cg[b + 3] * Q[(size_t)(b + 3) * n + i];
          }
          for (; b <= j; ++b) s0 += cg[b] * Q[(size_t)b * n + i];
          z[i] -= (s0 + s1) + (s2 + s3);
        }
        part = 0.0;
        for (int i = tid; i < n; i += blockDim.x) part += z[i] * z[i];
        const double after = block_sum(part, red);
        const bool again = after < 0.5 * zn2;
        zn2 = after;
        if (!again) break;
      }
      const double bj = sqrt(zn2);
      if (tid == 0) be[j] = bj;
      // breakdown: the Krylov space is invariant to working precision (the
      // Ritz pairs are exact); continuing would amplify rounding noise by 1/beta
      tmax = fmax(tmax, fabs(aj) + bj);
      const bool breakdown = !(bj > 64.0 * 2.220446049250313e-16 * tmax) || bj <= 1e-300;
      const double inv = breakdown ? 0.0 : 1.0 / bj;
      for (int i = tid; i < n; i += blockDim.x) {
        const double v = z[i] * inv;
        rq[i].y = v;
        Q[(size_t)(j + 1) * n + i] = v;
      }
      __syncthreads();
      const int m = j + 1;
      const bool check = m >= next_check || m == mmax || breakdown;
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
      // Ritz values to full precision, eigenvectors of T by twisted
      // factorisations (pivot workspace in LDS, batches of eigenpairs), exact
      // Ritz residuals |beta_m s_m|
      lz_multisect(al, e2, m, pivmin, ncand, nlow, hs, 2.0 * 2.220446049250313e-16, true, lo0,
                   hi0, clo, chi, cval);
      lz_select(cval, ncand, nlow, hs, kk, sval, sidx, cidx);
      {
        const int per = 2 * m;
        const int B = ivw / per > 0 ? ivw / per : 1;
        for (int c0 = 0; c0 < kk; c0 += B) {
          const int c = c0 + tid;
          if (tid < B && c < kk) {
            double* wsl = ivl + (size_t)tid * per;
            const double last = lz_twisted(al, be, m, sval[c], pivmin, wsl, wsl + m, nullptr);
            sres[c] = fabs(bj * last);
          }
          __syncthreads();
        }
      }
      const double tnorm = fmax(fabs(lo0), fabs(hi0));
      if (tid == 0) {
        double worst = 0.0;
        for (int c = 0; c < kk; ++c) worst = fmax(worst, sres[c]);
        const int conv = kk == k && worst <= tol * tnorm;
        ictl[0] = conv || m == mmax || breakdown;
        ictl[1] = kk;
        // status: 0 converged, 1 Krylov space exhausted (exact), 2 step cap hit
        ictl[2] = conv ? 0 : (breakdown ? 1 : 2);
        red[12] = worst;
        // next check from the geometric decay of the worst residual
        const double ratio = worst / (tol * tnorm);
        int gap = 5;
        if (prev_m > 0 && prev_worst > worst && worst > 0.0 && ratio > 1.0) {
          const double rate = log(prev_worst / worst) / (double)(m - prev_m);
          gap = (int)ceil(log(ratio) / rate);
          gap = gap < 2 ? 2 : (gap > 25 ? 25 : gap);
        }
        ictl[3] = m + gap;
      }
      __syncthreads();
      if (ictl[0]) {
        // final Ritz coefficient vectors into S
        const int per = 2 * m;
        const int B = ivw / per > 0 ? ivw / per : 1;
        for (int c0 = 0; c0 < kk; c0 += B) {
          const int c = c0 + tid;
          if (tid < B && c < kk) {
            double* wsl = ivl + (size_t)tid * per;
            lz_twisted(al, be, m, sval[c], pivmin, wsl, wsl + m, S + (size_t)c * mmax);
          }
          __syncthreads();
        }
      }
      if (ictl[0]) { m_final = m; break; }
      prev_m = m;
      prev_worst = red[12];
      next_check = ictl[3];
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
  return (size_t)(3 * nl + 4 * ml + 16 + 4 * kMaxNumCol + 4 * kMaxNumCol + lanczos_ivw(mmax)) *
             sizeof(double) +
         (size_t)(3 * kMaxNumCol + 8) * sizeof(int);
}

long long dcov_lanczos_scratch_doubles(int n, int k, int mmax) {
  long long s = (long long)n * (mmax + 1) + (long long)k * mmax + 64;
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
