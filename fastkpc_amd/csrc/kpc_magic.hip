// regrXonS design geometry and mgcv `magic` GCV smoothing selection on gfx950.
//
// geometry_kernel (one workgroup per conditioning set S):
//   X = [1 | X_1 | ... ] (fastkpc/src/full_cuda_ci_native_setup.cpp:1188-1317),
//   LAPACK-dgeqp3-style pivoted Householder QR of X, per-block penalty roots
//   by pivoted Cholesky with tol 0 (mgcv mroot "chol"), roots permuted into QR
//   order, S_i <- rS_i rS_i', and mgcv initial.sp
//   (full_cuda_ci_native_setup.cpp:881-1127; mgcv_multi_penalty_cpp.cpp:612-700).
// magic_kernel (one workgroup per (target, S) fit):
//   GCV score of [R; chol(sum_i e^rho_i S_i)] by one-sided Jacobi SVD with
//   rank tolerance sqrt(eps) (mgcv_multi_penalty_cpp.cpp:721-818), gradient and
//   Hessian in log(sp) (:819-990, evaluated through the factor T_i =
//   D^-1 V1' rS_i so every product is nr x rank_i), Newton / steepest descent
//   with step halving, the mgcv convergence test and the final infinity
//   probes (:1118-1346), then residuals y - X beta.
#include <algorithm>
#include <cstdlib>

#include "kpc_common.h"
#include "kpc_device.h"
#include "kpc_wave_linalg.h"

namespace kpc {

constexpr int kGeomThreads = 256;
// GCV fits run one per workgroup: 64 lanes for q <= 16 (level-1 designs),
// 256 lanes otherwise, so a fit's Jacobi rounds are spread over the four
// SIMDs of a CU (a wave64 instruction occupies one SIMD for four cycles and
// the optimiser's evaluations are strictly sequential).
constexpr int kMagicThreads = 256;
inline int magic_threads(int q) { return q <= 16 ? 64 : 256; }
constexpr double kEps = 2.220446049250313e-16;

__host__ __device__ inline DesignLayout design_layout(int n, int q, int m, int rk) {
  DesignLayout L;
  long long o = 0;
  L.X = o; o += (long long)n * q;
  L.QR = o; o += (long long)n * q;
  L.tau = o; o += q;
  L.R = o; o += (long long)q * q;
  L.piv = o; o += q;
  L.rS = o; o += (long long)q * m * rk;
  L.lsp0 = o; o += m;
  L.norms = o; o += 2 + 3 * m;
  L.status = o; o += 2;
  L.stride = (o + 31) / 32 * 32;
  return L;
}

DesignLayout host_design_layout(int n, int q, int m, int rk) { return design_layout(n, q, m, rk); }

// R mean of a short vector (serial; long double + correction in R).
__device__ __attribute__((always_inline)) double serial_r_mean(const double* x, int n) {
  dd s{0.0, 0.0};
  for (int i = 0; i < n; ++i) s = dd_add(s, dd{x[i], 0.0});
  dd m = dd_div_int(s, (double)n);
  dd c{0.0, 0.0};
  for (int i = 0; i < n; ++i) c = dd_add(c, dd_add(dd{x[i], 0.0}, dd{-m.hi, -m.lo}));
  dd r = dd_add(m, dd_div_int(c, (double)n));
  return r.hi + r.lo;
}

// Workgroup argmax with LAPACK MAXLOC semantics (first index among equal
// maxima); every thread receives the winning (value, index).
__device__ __attribute__((always_inline)) void block_argmax_first(double v, int idx, double* redv, int* redi, double* outv,
                                   int* outi) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int wid = threadIdx.x / KPC_WAVE;
  const int nw = (blockDim.x + KPC_WAVE - 1) / KPC_WAVE;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off, KPC_WAVE);
    const int oi = __shfl_xor(idx, off, KPC_WAVE);
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
  if (nw == 1) { *outv = v; *outi = idx; return; }
  __syncthreads();
  if (lane == 0) { redv[wid] = v; redi[wid] = idx; }
  __syncthreads();
  double bv = redv[0];
  int bi = redi[0];
  for (int w = 1; w < nw; ++w)
    if (redv[w] > bv || (redv[w] == bv && redi[w] < bi)) { bv = redv[w]; bi = redi[w]; }
  *outv = bv;
  *outi = bi;
}

// Pivoted Cholesky (LAPACK dpstf2, upper) of the q x q symmetric matrix A
// (column-major, ld q).  work: 2q doubles.  Returns the rank; piv receives the
// pivot order.  Every step is workgroup-parallel: diagonal updates, the
// first-maximum pivot search, the three disjoint row/column swaps and the
// row of R.  ictl needs 8 ints, dctl 8 doubles (plus redv/redi scratch of
// 16 entries each inside them).
__device__ __attribute__((always_inline)) int pstf2_upper(double* A, int q, int* piv, double* work, double tol, int* ictl,
                           double* dctl) {
  const int tid = threadIdx.x;
  double* redv = dctl + 8;
  int* redi = ictl + 8;
  for (int i = tid; i < q; i += blockDim.x) { piv[i] = i; work[i] = 0.0; }
  double bv = -__builtin_inf();
  int bi = 0x7fffffff;
  for (int i = tid; i < q; i += blockDim.x) {
    const double v = A[(size_t)i * q + i];
    if (v > bv || (v == bv && i < bi)) { bv = v; bi = i; }
  }
  double ajj0;
  int pvt0;
  block_argmax_first(bv, bi, redv, redi, &ajj0, &pvt0);
  if (ajj0 <= 0.0 || isnan(ajj0)) { __syncthreads(); return 0; }
  const double dstop = tol < 0.0 ? (double)q * kEps * ajj0 : tol;
  __syncthreads();
  int rank = q;
  for (int j = 0; j < q; ++j) {
    bv = -__builtin_inf();
    bi = 0x7fffffff;
    for (int i = j + tid; i < q; i += blockDim.x) {
      if (j > 0) work[i] += A[(size_t)i * q + j - 1] * A[(size_t)i * q + j - 1];
      const double w2 = A[(size_t)i * q + i] - work[i];
      work[q + i] = w2;
      if (w2 > bv || (w2 == bv && i < bi)) { bv = w2; bi = i; }
    }
    double ajj;
    int pvt;
    if (j == 0) {
      ajj = ajj0;
      pvt = pvt0;
    } else {
      block_argmax_first(bv, bi, redv, redi, &ajj, &pvt);
      if (ajj <= dstop || isnan(ajj)) {
        __syncthreads();
        if (tid == 0) A[(size_t)j * q + j] = ajj;
        rank = j;
        break;
      }
    }
    __syncthreads();
    if (j != pvt) {
      // the three swaps touch disjoint entries
      for (int k = tid; k < j; k += blockDim.x) {
        const double t = A[(size_t)j * q + k];
        A[(size_t)j * q + k] = A[(size_t)pvt * q + k];
        A[(size_t)pvt * q + k] = t;
      }
      for (int k = pvt + 1 + tid; k < q; k += blockDim.x) {
        const double t = A[(size_t)k * q + j];
        A[(size_t)k * q + j] = A[(size_t)k * q + pvt];
        A[(size_t)k * q + pvt] = t;
      }
      for (int k = j + 1 + tid; k < pvt; k += blockDim.x) {
        const double t = A[(size_t)k * q + j];
        A[(size_t)k * q + j] = A[(size_t)pvt * q + k];
        A[(size_t)pvt * q + k] = t;
      }
      if (tid == 0) {
        A[(size_t)pvt * q + pvt] = A[(size_t)j * q + j];
        const double t = work[j];
        work[j] = work[pvt];
        work[pvt] = t;
        const int it = piv[pvt];
        piv[pvt] = piv[j];
        piv[j] = it;
      }
      __syncthreads();
    }
    const double rjj = sqrt(ajj);
    const double rinv = 1.0 / rjj;
    for (int c = j + 1 + tid; c < q; c += blockDim.x) {
      double s = 0.0;
      for (int k = 0; k < j; ++k) s += A[(size_t)c * q + k] * A[(size_t)j * q + k];
      A[(size_t)c * q + j] = (A[(size_t)c * q + j] - s) * rinv;
    }
    __syncthreads();
    if (tid == 0) A[(size_t)j * q + j] = rjj;
    __syncthreads();
  }
  __syncthreads();
  return rank;
}

// X = R^-1 for the upper-triangular q x q matrix R (column-major, ld ld), X
// column-major with ld q, dinv: q doubles of scratch.  Right-looking: for
// k = q-1 .. 0 every (row i < k, column c >= k) pair subtracts R_ik x_kc in
// parallel, one workgroup barrier per k (the serial back substitution would
// chain q^2/2 dependent FP64 operations).  Rows are scaled by 1/R_ii at the end.
__device__ __attribute__((always_inline)) void upper_tri_inverse(const double* R, int ld, int q, double* X, double* dinv) {
  for (int e = threadIdx.x; e < q * q; e += blockDim.x) X[e] = (e % q == e / q) ? 1.0 : 0.0;
  for (int i = threadIdx.x; i < q; i += blockDim.x) dinv[i] = 1.0 / R[(size_t)i * ld + i];
  __syncthreads();
  for (int k = q - 1; k > 0; --k) {
    const double dk = dinv[k];
    const int w = q - k;
    for (int e = threadIdx.x; e < k * w; e += blockDim.x) {
      const int i = e % k, c = k + e / k;
      X[(size_t)c * q + i] -= R[(size_t)k * ld + i] * (X[(size_t)c * q + k] * dk);
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < q * q; e += blockDim.x) {
    const int i = e % q, c = e / q;
    if (i <= c) X[e] *= dinv[i];
  }
  __syncthreads();
}

typedef double v4d_t __attribute__((ext_vector_type(4)));

// C (M x N, column-major, ldc) = A' B for A (K x M, lda) and B (K x N, ldb),
// column-major, on FP64 matrix cores: the 16 x 16 output tiles are spread
// over the waves, each accumulating v_mfma_f64_16x16x4f64 over K in steps of
// 4 (lane (g, c) feeds A[k0 + g][i0 + c] and B[k0 + g][j0 + c]; entries past
// K, M or N read as zero).  No barrier inside: the caller orders phases.
__device__ __attribute__((always_inline)) void wg_gemm_tn(const double* A, int lda, const double* B, int ldb, double* C, int ldc,
                           int M, int N, int K) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int wid = threadIdx.x / KPC_WAVE;
  const int nw = blockDim.x / KPC_WAVE;
  const int g = lane >> 4, c = lane & 15;
  const int tm = (M + 15) / 16, tn = (N + 15) / 16;
  for (int t = wid; t < tm * tn; t += nw) {
    const int i0 = (t % tm) * 16, j0 = (t / tm) * 16;
    const int ia = i0 + c, jb = j0 + c;
    const double* a = A + (size_t)(ia < M ? ia : 0) * lda;
    const double* b = B + (size_t)(jb < N ? jb : 0) * ldb;
    v4d_t acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += 4) {
      const int k = k0 + g;
      const int kk = k < K ? k : 0;
      const double av = a[kk], bv = b[kk];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64((k < K && ia < M) ? av : 0.0,
                                                (k < K && jb < N) ? bv : 0.0, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = i0 + g + 4 * v, j = j0 + c;
      if (i < M && j < N) C[(size_t)j * ldc + i] = acc[v];
    }
  }
}

// ------------------------------------------------------------- geometry
__device__ __attribute__((always_inline)) void upper_tri_inverse_fast(const double* R, int ld, int q,
                                                                      double* X, double* dinv);

// kLdsQR: the pivoted QR works on an LDS copy of X (n x q, dynamic LDS) and
// stores the factored columns to the design afterwards; otherwise it works on
// the design's QR block in device memory.  Same operations either way.
template <bool kLdsQR>
__global__ __launch_bounds__(kGeomThreads) void geometry_kernel(
    const GeometryItem* __restrict__ items, int count, int n, int q, int m, int dim, int rk,
    const double* __restrict__ smooth_X, const double* __restrict__ smooth_S,
    double* __restrict__ design, double* __restrict__ status_out, long long smooth_x_stride,
    double* __restrict__ scratch, long long scratch_stride) {
  __shared__ double red[16];
  __shared__ double vn1[kMaxQ], vn2[kMaxQ], colv[kMaxQ];
  __shared__ int ipiv[kMaxQ];
  __shared__ int ictl[24];
  __shared__ double dctl[24];
  // initial.sp work of the serial section (thread 0), in LDS rather than
  // per-lane scratch memory
  __shared__ double isp_pd[kMaxQ], isp_vals[kMaxQ], isp_sp[kMaxPen], isp_sx[32], isp_spd[32];
  __shared__ unsigned char isp_pen[kMaxQ];
  __shared__ double isp_B[29 * 29];  // one smooth's penalty, staged for the serial section
  __shared__ double isp_rmn[29], isp_cmn[29];  // its row / column means of |B|
  // one smooth's penalty factor (dim <= 29), pivoted-Cholesky work and pivots
  __shared__ double pen_F[29 * 29], pen_wk[2 * 29 + 8];
  __shared__ int pen_pv[29];
  const int tid = threadIdx.x;
  const int lane = tid & (KPC_WAVE - 1);
  const int wid = tid / KPC_WAVE;
  const int nw = blockDim.x / KPC_WAVE;
  const DesignLayout DL = design_layout(n, q, m, rk);
  double* ws = scratch + (size_t)blockIdx.x * scratch_stride;
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const GeometryItem it = items[item];
    double* D = design + (size_t)it.design_slot * DL.stride;
    double* X = D + DL.X;
    extern __shared__ double geo_dyn[];
    double* QR = kLdsQR ? geo_dyn : D + DL.QR;
    // assemble X
    for (int r = tid; r < n; r += blockDim.x) X[r] = 1.0;
    for (int s = 0; s < m; ++s) {
      const double* Xs = smooth_X + (size_t)it.smooth_slot[s] * smooth_x_stride;
      for (int e = tid; e < n * dim; e += blockDim.x) X[(size_t)(1 + s * dim) * n + e] = Xs[e];
    }
    __syncthreads();
    for (int e = tid; e < n * q; e += blockDim.x) QR[e] = X[e];
    __syncthreads();
    // pivoted QR (dgeqp3 semantics)
    for (int j = wid; j < q; j += nw) {
      double s = 0.0;
      for (int r = lane; r < n; r += KPC_WAVE) s += QR[(size_t)j * n + r] * QR[(size_t)j * n + r];
      s = wave_sum(s);
      if (lane == 0) { vn1[j] = sqrt(s); vn2[j] = sqrt(s); ipiv[j] = j; }
    }
    __syncthreads();
    const double tol3z = sqrt(kEps);
    for (int i = 0; i < q; ++i) {
      if (tid == 0) {
        int pvt = i;
        for (int j = i + 1; j < q; ++j) if (vn1[j] > vn1[pvt]) pvt = j;
        ictl[0] = pvt;
        if (pvt != i) {
          int t = ipiv[pvt]; ipiv[pvt] = ipiv[i]; ipiv[i] = t;
          vn1[pvt] = vn1[i];
          vn2[pvt] = vn2[i];
        }
      }
      __syncthreads();
      const int pvt = ictl[0];
      if (pvt != i) {
        for (int r = tid; r < n; r += blockDim.x) {
          double t = QR[(size_t)pvt * n + r];
          QR[(size_t)pvt * n + r] = QR[(size_t)i * n + r];
          QR[(size_t)i * n + r] = t;
        }
        __syncthreads();
      }
      // larfg on column i rows i..n-1
      double part = 0.0;
      for (int r = i + 1 + tid; r < n; r += blockDim.x) part += QR[(size_t)i * n + r] * QR[(size_t)i * n + r];
      const double xn2 = block_sum(part, red);
      const double alpha = QR[(size_t)i * n + i];
      double tau = 0.0, beta = alpha, scal = 0.0;
      if (xn2 != 0.0) {
        beta = -copysign(hypot(alpha, sqrt(xn2)), alpha);
        tau = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
      }
      for (int r = i + 1 + tid; r < n; r += blockDim.x) QR[(size_t)i * n + r] *= scal;
      __syncthreads();
      if (tid == 0) {
        QR[(size_t)i * n + i] = beta;
        D[DL.tau + i] = tau;
      }
      __syncthreads();
      // apply H_i to trailing columns, a wave's columns j and j + nw side by
      // side (each column's sums in the same order as alone)
      if (tau != 0.0) {
        const double* vi = QR + (size_t)i * n;
        int j = i + 1 + wid;
        for (; j + nw < q; j += 2 * nw) {
          double* c0 = QR + (size_t)j * n;
          double* c1 = QR + (size_t)(j + nw) * n;
          double s0 = 0.0, s1 = 0.0;
          for (int r = i + lane; r < n; r += KPC_WAVE) {
            const double v = r == i ? 1.0 : vi[r];
            s0 += v * c0[r];
            s1 += v * c1[r];
          }
          s0 = wave_sum(s0) * tau;
          s1 = wave_sum(s1) * tau;
          for (int r = i + lane; r < n; r += KPC_WAVE) {
            const double v = r == i ? 1.0 : vi[r];
            c0[r] -= s0 * v;
            c1[r] -= s1 * v;
          }
        }
        if (j < q) {
          double* cj = QR + (size_t)j * n;
          double s = 0.0;
          for (int r = i + lane; r < n; r += KPC_WAVE) s += (r == i ? 1.0 : vi[r]) * cj[r];
          s = wave_sum(s) * tau;
          for (int r = i + lane; r < n; r += KPC_WAVE) cj[r] -= s * (r == i ? 1.0 : vi[r]);
        }
      }
      __syncthreads();
      // norm downdates
      for (int j = i + 1 + wid; j < q; j += nw) {
        if (vn1[j] != 0.0) {
          double temp = fabs(QR[(size_t)j * n + i]) / vn1[j];
          temp = 1.0 - temp * temp;
          if (temp < 0.0) temp = 0.0;
          double temp2 = temp * (vn1[j] / vn2[j]) * (vn1[j] / vn2[j]);
          if (temp2 <= tol3z) {
            double s = 0.0;
            for (int r = i + 1 + lane; r < n; r += KPC_WAVE) s += QR[(size_t)j * n + r] * QR[(size_t)j * n + r];
            s = wave_sum(s);
            if (lane == 0) { vn1[j] = sqrt(s); vn2[j] = vn1[j]; }
          } else if (lane == 0) {
            vn1[j] *= sqrt(temp);
          }
        }
      }
      __syncthreads();
    }
    // R and pivots (and the factored columns, from the LDS copy)
    if (kLdsQR)
      for (int e = tid; e < n * q; e += blockDim.x) D[DL.QR + e] = QR[e];
    for (int e = tid; e < q * q; e += blockDim.x) {
      const int r = e % q, c = e / q;
      D[DL.R + e] = r <= c ? QR[(size_t)c * n + r] : 0.0;
    }
    for (int i = tid; i < q; i += blockDim.x) D[DL.piv + i] = (double)ipiv[i];
    __syncthreads();
    // penalty roots: per block pivoted Cholesky with tol 0 (mroot)
    int status = 0;
    for (int s = 0; s < m; ++s) {
      const double* Sb = smooth_S + (size_t)it.smooth_slot[s] * 29 * 29;
      double* F = pen_F;   // dim x dim
      double* wk = pen_wk;  // 2 dim
      int* pv = pen_pv;
      for (int e = tid; e < dim * dim; e += blockDim.x) {
        const int r = e % dim, c = e / dim;
        F[e] = r <= c ? Sb[(size_t)c * dim + r] : 0.0;
      }
      __syncthreads();
      const int ar = pstf2_upper(F, dim, pv, wk, 0.0, ictl, dctl);
      if (ar < rk) status = -5;
      if (ar < dim) {
        for (int e = tid; e < dim * dim; e += blockDim.x) {
          const int r = e % dim, c = e / dim;
          if (c >= ar && r >= ar) F[e] = 0.0;
        }
      }
      __syncthreads();
      // inverse pivot -> block root (dim x rk): root[o][c] = F[c][inv[o]]
      double* rS = D + DL.rS + (size_t)s * q * rk;
      for (int e = tid; e < q * rk; e += blockDim.x) rS[e] = 0.0;
      __syncthreads();
      for (int e = tid; e < dim * rk; e += blockDim.x) {
        const int pos = e % dim, c = e / dim;  // pos = pivoted position
        const int orig = pv[pos];
        // full row index = 1 + s*dim + orig; then permute rows into QR order
        const int full_row = 1 + s * dim + orig;
        // find pivoted row r with ipiv[r] == full_row
        int prow = -1;
        for (int r = 0; r < q; ++r) if (ipiv[r] == full_row) { prow = r; break; }
        rS[(size_t)c * q + prow] = F[(size_t)pos * dim + c];
      }
      __syncthreads();
      // largest diagonal of S_i = rS_i rS_i' and its smallest kept Cholesky
      // pivot (the pivots of chol(sp_i S_i) are sp_i times these): the GCV
      // evaluation uses them to tell when the total-penalty root would drop
      // pivots
      double lm = 0.0;
      for (int r = tid; r < q; r += blockDim.x) {
        double v = 0.0;
        for (int k2 = 0; k2 < rk; ++k2) v += rS[(size_t)k2 * q + r] * rS[(size_t)k2 * q + r];
        lm = fmax(lm, v);
      }
      const double mxd = block_max(lm, red);
      if (tid == 0) {
        const double f = F[(size_t)(rk - 1) * dim + (rk - 1)];
        D[DL.norms + 2 + m + s] = mxd;
        D[DL.norms + 2 + 2 * m + s] = f * f;
      }
      __syncthreads();
    }
    // initial.sp
    for (int c = wid; c < q; c += nw) {
      dd s{0.0, 0.0};
      for (int r = lane; r < n; r += KPC_WAVE) {
        const double v = X[(size_t)c * n + r];
        s = dd_add(s, dd{v * v, 0.0});
      }
      s = wave_sum_dd(s);
      if (lane == 0) colv[c] = s.hi + s.lo;
    }
    __syncthreads();
    double* pd = isp_pd;  // q
    unsigned char* pen = isp_pen;
    double* sp = isp_sp;
    if (tid == 0)
      for (int c = 0; c < q; ++c) { pd[c] = 0.0; pen[c] = 0; }
    const double thr_scale = pow(kEps, 0.8);
    double* sx = isp_sx;
    double* sp_d = isp_spd;
    // per smooth: the workgroup stages its penalty in LDS, then thread 0 runs
    // the serial part from there (the same operations as reading it directly)
    for (int s = 0; s < m; ++s) {
      const double* Bg = smooth_S + (size_t)it.smooth_slot[s] * 29 * 29;
      __syncthreads();  // colv ready; the previous smooth's serial part is done
      for (int e = tid; e < dim * dim; e += blockDim.x) isp_B[e] = Bg[e];
      __syncthreads();
      // row and column means of |B| in double-double, a thread per row (each
      // sum in column order), and max |B| (exact in any order)
      for (int r = tid; r < dim; r += blockDim.x) {
        dd rt{0.0, 0.0}, ct{0.0, 0.0};
        for (int o = 0; o < dim; ++o) {
          rt = dd_add(rt, dd{fabs(isp_B[(size_t)o * dim + r]), 0.0});
          ct = dd_add(ct, dd{fabs(isp_B[(size_t)r * dim + o]), 0.0});
        }
        const dd rm = dd_div_int(rt, (double)dim), cmn = dd_div_int(ct, (double)dim);
        isp_rmn[r] = rm.hi + rm.lo;
        isp_cmn[r] = cmn.hi + cmn.lo;
      }
      double mxp = 0.0;
      for (int e = tid; e < dim * dim; e += blockDim.x) mxp = fmax(mxp, fabs(isp_B[e]));
      const double mx = block_max(mxp, red);
      if (tid == 0) {
        const double* B = isp_B;
        const int off = 1 + s * dim;
        const double thr = thr_scale * mx;
        int cnt = 0;
        for (int r = 0; r < dim; ++r) {
          const double rmn = isp_rmn[r], cmv2 = isp_cmn[r];
          const double dg = fabs(B[(size_t)r * dim + r]);
          if (rmn > thr && cmv2 > thr && dg > thr) {
            sx[cnt] = colv[off + r];
            sp_d[cnt] = B[(size_t)r * dim + r];
            pen[off + r] = 1;
            ++cnt;
          }
        }
        if (cnt == 0) {
          status = -8;
          sp[s] = 1.0;
        } else {
          sp[s] = serial_r_mean(sx, cnt) / serial_r_mean(sp_d, cnt);
          for (int r = 0; r < dim; ++r) pd[off + r] += sp[s] * B[(size_t)r * dim + r];
        }
      }
    }
    if (tid == 0) {
      double* vals = isp_vals;
      for (int guard = 0; guard < 4000; ++guard) {
        int cnt = 0;
        for (int c = 0; c < q; ++c)
          if (pd[c] > 0.0 && pen[c] && colv[c] > 0.0) vals[cnt++] = colv[c] / (colv[c] + pd[c]);
        if (serial_r_mean(vals, cnt) > 0.4) {
          for (int s = 0; s < m; ++s) sp[s] *= 10.0;
          for (int c = 0; c < q; ++c) pd[c] *= 10.0;
        } else {
          break;
        }
      }
      for (int guard = 0; guard < 4000; ++guard) {
        int cnt = 0;
        for (int c = 0; c < q; ++c)
          if (pd[c] > 0.0 && pen[c] && colv[c] > 0.0) vals[cnt++] = colv[c] / (colv[c] + pd[c]);
        if (serial_r_mean(vals, cnt) < 0.4) {
          for (int s = 0; s < m; ++s) sp[s] /= 10.0;
          for (int c = 0; c < q; ++c) pd[c] /= 10.0;
        } else {
          break;
        }
      }
      for (int s = 0; s < m; ++s) D[DL.lsp0 + s] = log(sp[s]);
      D[DL.status] = (double)status;
      if (status_out) status_out[it.design_slot] = (double)status;
    }
    __syncthreads();
    // norms for the GCV full-rank certificate: |R|_F^2, |R^-1|_F^2, |rS_i|_F^2
    {
      double* Rinv = ws + q;  // q x q after the q doubles of initial.sp
      upper_tri_inverse_fast(D + DL.R, q, q, Rinv, Rinv + (size_t)q * q);
      double pr = 0.0, pi = 0.0;
      for (int e = tid; e < q * q; e += blockDim.x) {
        const int r = e % q, c = e / q;
        if (r <= c) {
          const double a = D[DL.R + e];
          pr += a * a;
          pi += Rinv[e] * Rinv[e];
        }
      }
      const double nr2 = block_sum(pr, red);
      const double ni2 = block_sum(pi, red);
      if (tid == 0) {
        D[DL.norms] = nr2;
        D[DL.norms + 1] = ni2;
      }
      for (int i = 0; i < m; ++i) {
        double ps = 0.0;
        for (int e = tid; e < rk * q; e += blockDim.x) {
          const double v = D[DL.rS + (size_t)i * rk * q + e];
          ps += v * v;
        }
        const double ns = block_sum(ps, red);
        if (tid == 0) D[DL.norms + 2 + i] = ns;
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------- magic
// Per-fit working set.  The big arrays (aug, V, X4, Qm: 6 q^2 + 3 q) and the
// derivative work (3 m q + 4 m^2 + 3 m + 8) live in LDS when they fit and in
// the fit's global slab otherwise; every design here has m rk <= q (tp: rank
// q - 2; additive: 8 m < 1 + 9 m).
__host__ __device__ inline long long magic_big_doubles(int q) {
  return 6LL * q * q + 3LL * q;
}
__host__ __device__ inline long long magic_deriv_doubles(int q, int m) {
  return 3LL * m * q + 4LL * m * m + 3LL * m + 8;
}
__host__ __device__ inline long long magic_small_doubles(int q, int m) {
  // d, y1, wk (4q), 7 m-vectors + hessian, red, dctl, spv (2m), design norms
  // (2 + 3m), ord/piv/ictl
  return 4LL * q + 7LL * m + (long long)m * m + 16 + 32 + 2LL * m + (2 + 3LL * m) + q + 24;
}
__host__ __device__ inline bool magic_deriv_in_lds(int q, int m) {
  return magic_deriv_doubles(q, m) * 8 <= 48 * 1024;
}
// Doubles of the per-fit global slab ahead of the response vectors.
__host__ __device__ inline long long magic_scratch_core(int q, int m) {
  long long s = magic_big_doubles(q) + magic_deriv_doubles(q, m) + 2LL * q + 64;
  return (s + 31) / 32 * 32;
}

struct MagicLds {
  double* y;      // n (Q'y in the first q entries after the transform)
  double* yorig;  // n
  double* aug;    // (q + q) x q, column-major with ld = rows
  double* V;      // q x q: SVD right vectors, or R2^-1 on the QR path
  double* X4;     // q x q: aggregate-penalty factor, then G = U1'U1
  double* Qm;     // (q + q) x q: orthogonal factor of [R; E] on the QR path
  double* gyl;    // q: G y1
  double* projl;  // 2q: T_a' y1, T_a' G y1
  double* d;      // q singular values (by column)
  int* ord;       // q: descending singular value order
  double* y1;     // q
  double* lsp;    // m
  double* trial;  // m
  double* grad;   // m (gradient used by the convergence test)
  double* cgrad;  // m (gradient at the current point)
  double* hess;   // m^2
  double* nstep;
  double* sstep;
  double* step;
  double* red;    // 16
  double* dctl;   // 32
  int* ictl;      // 32
  int* piv;       // q
  double* wk;     // 2q
  double* spv;    // 2m: sp_i = exp(log sp_i), then sqrt(sp_i) at m + i
  double* nrm;    // 2 + 3m: the design's norms block, read once per fit
  double* My;     // m x nr derivative work
  double* Ky;     // m x nr
  double* yK;     // m x nr
  double* tr;     // m^2 + m: pairwise traces, then tr(K_i)
  double* dn;     // m: d|y - Ay|^2 / d rho_i
  double* d2n;    // m x m
  double* syA;    // m x m: Hessian eigen work
  double* syZ;    // m x m: eigenvectors
  double* syw;    // m: eigenvalues
  // trial slots (null unless the launch scores step-halving trials on
  // several waves at once): the fit's constant inputs staged once per fit
  // and each slot's own outputs
  double* Rst;     // q x q: the design's R (column-major)
  double* rSst;    // m rk x q: penalty roots rS
  double* yst;     // q: Q'y
  double* trials;  // kSpecTries x m: one iteration's trial points in try order
  double* sres;    // 5 x 8: per-point {score, rss, edf, rank margin, certified}
  double* sy1;     // 4 x q: y1 of the points scored outside the workgroup's buffers
  double* sQ3;     // q x q: Q1 of the third such point (the first two use X4 and V)
  double* sQ4;     // q x q: Q1 of the fourth (wave 0's second point; five slots)
  int nslot;
};

// Trial slots: a multi-penalty additive design (rank-8 penalty blocks,
// q = 1 + 9 m, m = 3 or 4) on a 256-lane workgroup whose working set is all
// in LDS scores up to nslot step-halving trials at once, one per wave, and at
// m = 3 a fifth on wave 0 beside its first (wave_eval_additive2).  (At
// m = 5 the one-wave evaluation holds 60 doubles of E and X per lane and
// measured twice as slow as the workgroup path on fits without halving
// tails, so m = 5 keeps magic_kernel.)
constexpr int kSpecTries = 32;
__host__ __device__ inline bool magic_slot_shape(int q, int m, int rk) {
  return rk == 8 && m >= 3 && m <= 4 && q == 1 + 9 * m;
}
__host__ __device__ inline long long magic_spec_doubles(int q, int m, int nslot) {
  if (nslot <= 1) return 0;
  return (long long)q * q + 8LL * m * q + q + (long long)kSpecTries * m + 40 + 4LL * q +
         (nslot >= 4 ? (long long)q * q : 0) + (nslot >= 5 ? (long long)q * q : 0);
}

__host__ __device__ inline size_t magic_lds_doubles(int n, int q, int m, bool aug_in_lds,
                                                    int nslot = 1) {
  (void)n;
  size_t s = (aug_in_lds ? (size_t)magic_big_doubles(q) : 0) + (size_t)magic_small_doubles(q, m);
  if (magic_deriv_in_lds(q, m)) s += (size_t)magic_deriv_doubles(q, m);
  return s + (size_t)magic_spec_doubles(q, m, nslot);
}

// Slots for (n, q, m, rk): 5 at m = 3 and 4 otherwise while the workgroup's
// LDS stays within half a CU (two fits per CU), else 3 (slots 1 and 2 reuse
// V and X4) within the single-fit budget, else 1 (trials scored one at a
// time).
inline int magic_spec_slots(int n, int q, int m, int rk) {
  if (!magic_slot_shape(q, m, rk) || q <= 16 || !magic_deriv_in_lds(q, m)) return 1;
  const size_t base = magic_lds_doubles(n, q, m, true);
  if (m == 3 && (base + (size_t)magic_spec_doubles(q, m, 5)) * 8 <= 80 * 1024) return 5;
  if ((base + (size_t)magic_spec_doubles(q, m, 4)) * 8 <= 80 * 1024) return 4;
  if ((base + (size_t)magic_spec_doubles(q, m, 3)) * 8 <= 150 * 1024) return 3;
  return 1;
}

__device__ __attribute__((always_inline)) MagicLds carve_magic(double* base, int n, int q, int m, bool aug_in_lds,
                                double* gscratch, bool deriv_in_lds, int nslot = 1) {
  MagicLds L;
  double* p = base;
  // response vectors live in the global slab after the working set
  L.y = gscratch + magic_scratch_core(q, m);
  L.yorig = L.y + n;
  double* big = aug_in_lds ? p : gscratch;
  L.aug = big;
  L.V = big + 2 * (size_t)q * q;
  L.X4 = big + 3 * (size_t)q * q;
  L.Qm = big + 4 * (size_t)q * q;
  L.gyl = big + 6 * (size_t)q * q;
  L.projl = L.gyl + q;
  if (aug_in_lds) p += magic_big_doubles(q);
  L.d = p; p += q;
  L.y1 = p; p += q;
  L.lsp = p; p += m;
  L.trial = p; p += m;
  L.grad = p; p += m;
  L.cgrad = p; p += m;
  L.nstep = p; p += m;
  L.sstep = p; p += m;
  L.step = p; p += m;
  L.hess = p; p += (size_t)m * m;
  L.red = p; p += 16;
  L.dctl = p; p += 32;
  L.wk = p; p += 2 * q;
  L.spv = p; p += 2 * m;
  L.nrm = p; p += 2 + 3 * m;
  double* dw = deriv_in_lds ? p : gscratch + magic_big_doubles(q);
  L.My = dw; dw += (size_t)m * q;
  L.Ky = dw; dw += (size_t)m * q;
  L.yK = dw; dw += (size_t)m * q;
  L.tr = dw; dw += (size_t)m * m + m;
  L.dn = dw; dw += m;
  L.d2n = dw; dw += (size_t)m * m;
  L.syA = dw; dw += (size_t)m * m;
  L.syZ = dw; dw += (size_t)m * m;
  L.syw = dw; dw += m + 8;
  if (deriv_in_lds) p = dw;
  int* ip = reinterpret_cast<int*>(p);
  L.ord = ip; ip += q;
  L.piv = ip; ip += q;
  L.ictl = ip;
  ip += 32;
  L.nslot = nslot > 1 && aug_in_lds && deriv_in_lds ? nslot : 1;
  L.Rst = L.rSst = L.yst = L.trials = L.sres = L.sy1 = L.sQ3 = L.sQ4 = nullptr;
  if (L.nslot > 1) {
    double* sp = reinterpret_cast<double*>(ip);  // 2 q + 32 ints: 8-byte aligned
    L.Rst = sp; sp += (size_t)q * q;
    L.rSst = sp; sp += (size_t)8 * m * q;
    L.yst = sp; sp += q;
    L.trials = sp; sp += (size_t)kSpecTries * m;
    L.sres = sp; sp += 40;
    L.sy1 = sp; sp += 4 * (size_t)q;
    L.sQ3 = L.nslot >= 4 ? sp : nullptr;
    if (L.nslot >= 4) sp += (size_t)q * q;
    L.sQ4 = L.nslot >= 5 ? sp : nullptr;
  }
  return L;
}

struct MagicScratch {
  double* theta;  // q: coefficients in the QR basis
  double* beta;   // q: coefficients in the original column order
};

__device__ __attribute__((always_inline)) MagicScratch carve_magic_scratch(double* base, int q, int m, int rk) {
  (void)rk;
  MagicScratch S;
  double* p = base + magic_big_doubles(q) + magic_deriv_doubles(q, m);
  S.theta = p; p += q;
  S.beta = p; p += q;
  return S;
}

long long magic_scratch_doubles(int q, int m, int rk, int n) {
  long long s = magic_scratch_core(q, m) + 2LL * n + 64;
  (void)m;
  (void)rk;
  return (s + 31) / 32 * 32;
}

// One-sided Jacobi SVD of the rows x q matrix A (column-major, ld rows) with
// round-robin parallel ordering; V accumulates the rotations.
// Value of lane (lane ^ OFF) within aligned groups of 2 OFF lanes: DPP moves
// for OFF < 16 (quad permutations, row shifts by 4 / 8 with a select), a
// cross-lane read otherwise.
template <int OFF>
__device__ __forceinline__ double xor_lanes(double v);

// Sweeps of jacobi_svd with the pair width lpp = LPP and at most R rows per
// lane (rows <= LPP R), every pair of a round in one pass: each lane keeps its
// rows of the pair's columns (and of V) in registers from the dot products to
// the rotation, the round-robin partners come from conditional subtractions,
// and the xor reductions are DPP moves.  Same operations in the same order as
// the general sweep loop (same rows per lane, same reduction tree, same
// rotation), so the results are bit-identical.
template <int LPP, int R>
__device__ __attribute__((always_inline)) void jacobi_sweeps_fixed(double* A, int rows, int q,
                                                                   double* V, int* ictl) {
  constexpr int kJc = 8;
  const int tid = threadIdx.x;
  const int qe = q + (q & 1);
  const int npairs = qe / 2;
  const int qm = qe - 1;
  const int pair = tid / LPP, sub = tid % LPP;
  for (int sweep = 0; sweep < 60; ++sweep) {
    if (tid == 0) ictl[8] = 0;
    __syncthreads();
    for (int rnd = 0; rnd < qm; ++rnd) {
      int a = -1, b = -1;
      if (pair < npairs) {
        if (pair == 0) {
          a = rnd;
          b = qe - 1;
        } else {
          a = rnd + pair;
          if (a >= qm) a -= qm;
          b = rnd - pair;
          if (b < 0) b += qm;
        }
        if (a > b) {
          const int t = a;
          a = b;
          b = t;
        }
      }
      const bool active = pair < npairs && b < q;
      double xs[R], ys[R], vx[R], vy[R];
      double al = 0.0, be = 0.0, ga = 0.0;
      if (active) {
        const double* ca = A + (size_t)a * rows;
        const double* cb = A + (size_t)b * rows;
        const double* va = V + (size_t)a * q;
        const double* vb = V + (size_t)b * q;
#pragma unroll
        for (int u = 0; u < R; ++u) {
          const int r = sub + u * LPP;
          xs[u] = r < rows ? ca[r] : 0.0;
          ys[u] = r < rows ? cb[r] : 0.0;
          vx[u] = r < q ? va[r] : 0.0;
          vy[u] = r < q ? vb[r] : 0.0;
        }
        // (the general loop runs chunks of kJc rows; with R <= kJc that is one
        // chunk plus rows past the column, which add +0)
#pragma unroll
        for (int u = 0; u < kJc; ++u) {
          const double x = u < R ? xs[u] : 0.0, y = u < R ? ys[u] : 0.0;
          al += x * x;
          be += y * y;
          ga += x * y;
        }
      }
#pragma unroll
      for (int off = LPP >> 1; off > 0; off >>= 1) {
        double ta, tb, tg;
        switch (off) {
          case 1: ta = xor_lanes<1>(al); tb = xor_lanes<1>(be); tg = xor_lanes<1>(ga); break;
          case 2: ta = xor_lanes<2>(al); tb = xor_lanes<2>(be); tg = xor_lanes<2>(ga); break;
          case 4: ta = xor_lanes<4>(al); tb = xor_lanes<4>(be); tg = xor_lanes<4>(ga); break;
          case 8: ta = xor_lanes<8>(al); tb = xor_lanes<8>(be); tg = xor_lanes<8>(ga); break;
          default: ta = xor_lanes<16>(al); tb = xor_lanes<16>(be); tg = xor_lanes<16>(ga); break;
        }
        al += ta;
        be += tb;
        ga += tg;
      }
      if (active && ga != 0.0 && fabs(ga) > kEps * sqrt(al * be)) {
        const double zeta = (be - al) / (2.0 * ga);
        const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t);
        const double sn = c * t;
        double* ca = A + (size_t)a * rows;
        double* cb = A + (size_t)b * rows;
        double* va = V + (size_t)a * q;
        double* vb = V + (size_t)b * q;
#pragma unroll
        for (int u = 0; u < R; ++u) {
          const int r = sub + u * LPP;
          if (r < rows) {
            ca[r] = c * xs[u] - sn * ys[u];
            cb[r] = sn * xs[u] + c * ys[u];
          }
          if (r < q) {
            va[r] = c * vx[u] - sn * vy[u];
            vb[r] = sn * vx[u] + c * vy[u];
          }
        }
        if (sub == 0) ictl[8] = 1;
      }
      __syncthreads();
    }
    if (tid == 0) ictl[28] += 1;
    if (!ictl[8]) break;
    __syncthreads();
  }
}

__device__ __attribute__((always_inline)) void jacobi_svd(double* A, int rows, int q, double* V, double* d, int* ord,
                           int* ictl) {
  const int tid = threadIdx.x;
  for (int e = tid; e < q * q; e += blockDim.x) V[e] = 0.0;
  __syncthreads();
  for (int i = tid; i < q; i += blockDim.x) V[(size_t)i * q + i] = 1.0;
  const int qe = q + (q & 1);
  const int npairs = qe / 2;
  int lpp = 64;
  while (lpp > 1 && npairs * lpp > (int)blockDim.x) lpp >>= 1;
  const int groups = (int)blockDim.x / lpp;
  const int sub = tid % lpp;
  const int qm = qe - 1;
  // pairs per group (more than one only for q > 2 blockDim), uniform count
  const int passes = (npairs + groups - 1) / groups;
  __syncthreads();
  const bool fixed8 = lpp == 8 && passes == 1 && rows <= 8 * 8;
  const bool fixed16 = lpp == 16 && passes == 1 && rows <= 16 * 4;
  if (fixed8) jacobi_sweeps_fixed<8, 8>(A, rows, q, V, ictl);
  if (fixed16) jacobi_sweeps_fixed<16, 4>(A, rows, q, V, ictl);
  for (int sweep = 0; sweep < 60 && !fixed8 && !fixed16; ++sweep) {
    if (tid == 0) ictl[8] = 0;
    __syncthreads();
    // round-robin schedule: pair 0 meets the fixed column qe-1, pair p>0
    // pairs (rnd + p) and (rnd - p) mod (qe - 1)
    for (int rnd = 0; rnd < qm; ++rnd) {
     for (int ps = 0; ps < passes; ++ps) {
      const int pair = tid / lpp + ps * groups;
      int a = -1, b = -1;
      if (pair < npairs) {
        if (pair == 0) { a = rnd; b = qe - 1; }
        else { a = (rnd + pair) % qm; b = (rnd - pair % qm + qm) % qm; }
        if (a > b) { int t = a; a = b; b = t; }
      }
      const bool active = pair < npairs && b < q;
      double al = 0.0, be = 0.0, ga = 0.0;
      // rows are visited in chunks of kJc with all loads of a chunk issued
      // before use (LDS latency overlaps); per-lane accumulation order is the
      // plain row order
      constexpr int kJc = 8;
      if (active) {
        const double* ca = A + (size_t)a * rows;
        const double* cb = A + (size_t)b * rows;
        for (int r0 = sub; r0 < rows; r0 += lpp * kJc) {
          double xs[kJc], ys[kJc];
#pragma unroll
          for (int u = 0; u < kJc; ++u) {
            const int r = r0 + u * lpp;
            xs[u] = r < rows ? ca[r] : 0.0;
            ys[u] = r < rows ? cb[r] : 0.0;
          }
#pragma unroll
          for (int u = 0; u < kJc; ++u) {
            al += xs[u] * xs[u];
            be += ys[u] * ys[u];
            ga += xs[u] * ys[u];
          }
        }
      }
      for (int off = lpp >> 1; off > 0; off >>= 1) {
        al += __shfl_xor(al, off, KPC_WAVE);
        be += __shfl_xor(be, off, KPC_WAVE);
        ga += __shfl_xor(ga, off, KPC_WAVE);
      }
      if (active && ga != 0.0 && fabs(ga) > kEps * sqrt(al * be)) {
        const double zeta = (be - al) / (2.0 * ga);
        const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t);
        const double sn = c * t;
        double* ca = A + (size_t)a * rows;
        double* cb = A + (size_t)b * rows;
        for (int r0 = sub; r0 < rows; r0 += lpp * kJc) {
          double xs[kJc], ys[kJc];
#pragma unroll
          for (int u = 0; u < kJc; ++u) {
            const int r = r0 + u * lpp;
            xs[u] = r < rows ? ca[r] : 0.0;
            ys[u] = r < rows ? cb[r] : 0.0;
          }
#pragma unroll
          for (int u = 0; u < kJc; ++u) {
            const int r = r0 + u * lpp;
            if (r < rows) {
              ca[r] = c * xs[u] - sn * ys[u];
              cb[r] = sn * xs[u] + c * ys[u];
            }
          }
        }
        double* va = V + (size_t)a * q;
        double* vb = V + (size_t)b * q;
        for (int r0 = sub; r0 < q; r0 += lpp * kJc) {
          double xs[kJc], ys[kJc];
#pragma unroll
          for (int u = 0; u < kJc; ++u) {
            const int r = r0 + u * lpp;
            xs[u] = r < q ? va[r] : 0.0;
            ys[u] = r < q ? vb[r] : 0.0;
          }
#pragma unroll
          for (int u = 0; u < kJc; ++u) {
            const int r = r0 + u * lpp;
            if (r < q) {
              va[r] = c * xs[u] - sn * ys[u];
              vb[r] = sn * xs[u] + c * ys[u];
            }
          }
        }
        if (sub == 0) ictl[8] = 1;
      }
     }
      __syncthreads();
    }
    if (tid == 0) ictl[28] += 1;
    if (!ictl[8]) break;
    __syncthreads();
  }
  // singular values and normalised left vectors
  const int lane = tid & (KPC_WAVE - 1);
  const int wid = tid / KPC_WAVE;
  const int nw = blockDim.x / KPC_WAVE;
  for (int c = wid; c < q; c += nw) {
    double* col = A + (size_t)c * rows;
    double s = 0.0;
    for (int r = lane; r < rows; r += KPC_WAVE) s += col[r] * col[r];
    s = sqrt(wave_sum(s));
    if (lane == 0) d[c] = s;
    const double inv = s > 0.0 ? 1.0 / s : 0.0;
    for (int r = lane; r < rows; r += KPC_WAVE) col[r] *= inv;
  }
  __syncthreads();
  // descending order, ties keep column order (stable): rank by counting
  for (int i = tid; i < q; i += blockDim.x) {
    const double di = d[i];
    int r = 0;
    for (int j = 0; j < q; ++j) r += (d[j] > di) || (d[j] == di && j < i);
    ord[r] = i;
  }
  __syncthreads();
}

// 8-lane sum (aligned groups of 8 lanes) by DPP moves: xor 1, xor 2 inside
// quads, then the half-row mirror.
template <int kCtrl>
__device__ __forceinline__ double dpp_move(double v) {
  // every lane of an 8-lane group is active whenever a group sums, so no
  // lane reads a disabled source and the moves need no fill value
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), kCtrl, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), kCtrl, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

template <int OFF>
__device__ __forceinline__ double xor_lanes(double v) {
  if constexpr (OFF == 1) {
    return dpp_move<0xB1>(v);  // quad_perm [1,0,3,2]
  } else if constexpr (OFF == 2) {
    return dpp_move<0x4E>(v);  // quad_perm [2,3,0,1]
  } else if constexpr (OFF == 4 || OFF == 8) {
    const double up = dpp_move<0x100 + OFF>(v);  // row_shl: lane + OFF
    const double dn = dpp_move<0x110 + OFF>(v);  // row_shr: lane - OFF
    return (threadIdx.x & OFF) ? dn : up;
  } else {
    return __shfl_xor(v, OFF, KPC_WAVE);
  }
}

__device__ __forceinline__ double sum8(double v) {
  v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x141>(v);  // row_half_mirror
  return v;
}

// Householder reflector (LAPACK dlarfg) of column jj of W over row jj and the
// E rows q..q+mr-1, by one aligned group of 8 lanes: v overwrites the E part,
// beta the diagonal, tau[jj] the scale.
__device__ __forceinline__ void group_reflector(double* W, int rows, int q, int mr, int jj,
                                                double* tau, int sub) {
  double* w = W + (size_t)jj * rows;
  double s2 = 0.0;
  for (int k = sub; k < mr; k += 8) s2 += w[q + k] * w[q + k];
  s2 = sum8(s2);
  wave_sync();
  const double alpha = w[jj];
  double t = 0.0, scal = 0.0, beta = alpha;
  if (s2 > 0.0) {
    // r = 1/|(alpha, x)|: beta = -sign(alpha)/r, tau = 1 + |alpha| r,
    // 1/(alpha - beta) = 1 / (sign(alpha) (|alpha| + 1/r))
    const double h = fma(alpha, alpha, s2);
    const double r = nrsq(h);
    const double nrm = h * r;
    beta = -copysign(nrm, alpha);
    t = fma(fabs(alpha), r, 1.0);
    scal = copysign(nrcp(fabs(alpha) + nrm), alpha);
  }
  for (int k = sub; k < mr; k += 8) w[q + k] *= scal;
  if (sub == 0) {
    w[jj] = beta;
    tau[jj] = t;
  }
}

// Structured Householder QR of W = [R; E] (rows = q + mr, column-major, ld
// rows; R upper triangular): step j reflects row j of the top block against
// the E rows only, so the top block stays triangular and every reflector has
// length 1 + mr.  Groups of 8 lanes own one column each; the group that
// updates column j + 1 builds its reflector right away, so each step costs
// one workgroup barrier.  R2 overwrites the top block, reflector tails the E
// rows, tau[0..q).
__device__ __attribute__((always_inline)) void stacked_qr(double* W, int rows, int q, int mr, double* tau) {
  const int tid = threadIdx.x;
  constexpr int kG = 8;
  const int ncg = blockDim.x / kG;
  const int sub = tid % kG, cg = tid / kG;
  if (cg == 0) group_reflector(W, rows, q, mr, 0, tau, sub);
  __syncthreads();
  for (int j = 0; j < q - 1; ++j) {
    const double tj = tau[j];
    const double* vj = W + (size_t)j * rows + q;
    for (int c0 = j + 1; c0 < q; c0 += ncg) {
      const int c = c0 + cg;
      double* wc = W + (size_t)(c < q ? c : j) * rows;
      double dot = 0.0;
      if (c < q) {
        if (sub == 0) dot = wc[j];
        for (int k = sub; k < mr; k += kG) dot += vj[k] * wc[q + k];
      }
      dot = sum8(dot) * tj;
      if (c < q && tj != 0.0) {
        if (sub == 0) wc[j] -= dot;
        for (int k = sub; k < mr; k += kG) wc[q + k] -= dot * vj[k];
      }
      if (c == j + 1) group_reflector(W, rows, q, mr, j + 1, tau, sub);
    }
    __syncthreads();
  }
}

// Q1 (top q rows of the orthogonal factor of [R; E]) from the reflectors of
// stacked_qr: Qm = H_0 ... H_{q-1} [I; 0] (LAPACK dorg2r order).
__device__ __attribute__((always_inline)) void stacked_q1(const double* W, int rows, int q, int mr, const double* tau,
                           double* Qm) {
  const int tid = threadIdx.x;
  constexpr int kG = 8;
  const int ncg = blockDim.x / kG;
  const int sub = tid % kG, cg = tid / kG;
  for (int e = tid; e < rows * q; e += blockDim.x) Qm[e] = (e % rows == e / rows) ? 1.0 : 0.0;
  __syncthreads();
  for (int j = q - 1; j >= 0; --j) {
    const double tj = tau[j];
    if (tj == 0.0) continue;
    const double* v = W + (size_t)j * rows + q;
    for (int c0 = j; c0 < q; c0 += ncg) {
      const int c = c0 + cg;
      double* qc = Qm + (size_t)(c < q ? c : j) * rows;
      double dot = 0.0;
      if (c < q) {
        if (sub == 0) dot = qc[j];
        for (int k = sub; k < mr; k += kG) dot += v[k] * qc[q + k];
      }
      dot = sum8(dot) * tj;
      if (c < q) {
        if (sub == 0) qc[j] -= dot;
        for (int k = sub; k < mr; k += kG) qc[q + k] -= dot * v[k];
      }
    }
    __syncthreads();
  }
}

// ---- register-resident forms (q, mr <= 8 K): the same arithmetic, in the
// same order per column, as stacked_qr / stacked_q1 / upper_tri_inverse, with
// each column held by one aligned group of 8 lanes (lane sub holds rows
// sub, sub + 8, ...).  Only the exchange of data changes, so results are
// bit-identical: Q1 and R^-1 need no workgroup barrier at all, the QR one per
// reflector instead of LDS round trips for every column entry.

// Q1 column by column: column c only ever combines with itself, so every
// group runs its columns (cg, cg + ng) through reflectors q-1 .. 0 alone.
// Needs q <= 2 ng.
template <int K>
__device__ __attribute__((always_inline)) void stacked_q1_regs(const double* W, int rows, int q,
                                                               int mr, const double* tau,
                                                               double* Qm) {
  const int tid = threadIdx.x;
  const int ng = (int)blockDim.x >> 3, cg = tid >> 3, sub = tid & 7;
  for (int e = tid; e < rows * q; e += blockDim.x) Qm[e] = (e % rows == e / rows) ? 1.0 : 0.0;
  __syncthreads();
  double E[2][K];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int t = 0; t < K; ++t) E[i][t] = 0.0;
  double vn[K];
  double tn = tau[q - 1];
#pragma unroll
  for (int t = 0; t < K; ++t)
    vn[t] = sub + 8 * t < mr ? W[(size_t)(q - 1) * rows + q + sub + 8 * t] : 0.0;
  for (int j = q - 1; j >= 0; --j) {
    double v[K];
#pragma unroll
    for (int t = 0; t < K; ++t) v[t] = vn[t];
    const double tj = tn;
    if (j > 0) {
      tn = tau[j - 1];
#pragma unroll
      for (int t = 0; t < K; ++t)
        vn[t] = sub + 8 * t < mr ? W[(size_t)(j - 1) * rows + q + sub + 8 * t] : 0.0;
    }
    if (tj == 0.0) continue;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int c = cg + ng * i;
      if (c < q && c >= j) {
        const double top = c == j ? 1.0 : 0.0;
        double dot = 0.0;
        if (sub == 0) dot = top;
#pragma unroll
        for (int t = 0; t < K; ++t)
          if (sub + 8 * t < mr) dot += v[t] * E[i][t];
        dot = sum8(dot) * tj;
        if (sub == 0) Qm[(size_t)c * rows + j] = top - dot;
#pragma unroll
        for (int t = 0; t < K; ++t)
          if (sub + 8 * t < mr) E[i][t] -= dot * v[t];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int c = cg + ng * i;
#pragma unroll
    for (int t = 0; t < K; ++t) {
      const int k = sub + 8 * t;
      if (c < q && k < mr) Qm[(size_t)c * rows + q + k] = E[i][t];
    }
  }
  __syncthreads();
}

// Structured QR with the columns spread over every 8-lane group of the
// workgroup (columns cg, cg + ng), their E rows in registers: reflector j is
// built by the group holding column j right after that column's update and
// published through W (tail), tau and the diagonal; one workgroup barrier per
// step.  Needs q <= 2 ng.
template <int K>
__device__ __attribute__((always_inline)) void qr_cols_reflector(double (&e)[K], double* W,
                                                                  int rows, int q, int mr, int jj,
                                                                  double* tau, int sub) {
  double s2 = 0.0;
#pragma unroll
  for (int t = 0; t < K; ++t)
    if (sub + 8 * t < mr) s2 += e[t] * e[t];
  s2 = sum8(s2);
  const double alpha = W[(size_t)jj * rows + jj];
  double tj = 0.0, scal = 0.0, beta = alpha;
  if (s2 > 0.0) {
    const double h = fma(alpha, alpha, s2);
    const double r = nrsq(h);
    const double nrm = h * r;
    beta = -copysign(nrm, alpha);
    tj = fma(fabs(alpha), r, 1.0);
    scal = copysign(nrcp(fabs(alpha) + nrm), alpha);
  }
#pragma unroll
  for (int t = 0; t < K; ++t) {
    const int k = sub + 8 * t;
    if (k < mr) {
      e[t] *= scal;
      W[(size_t)jj * rows + q + k] = e[t];
    }
  }
  if (sub == 0) {
    W[(size_t)jj * rows + jj] = beta;
    tau[jj] = tj;
  }
}

// Q1 alongside the QR: X = Q' [I; 0] = H_{q-1} ... H_0 [I; 0] takes every
// reflector as soon as it is published, and its top block is Q1' (the top
// entry of column c in row j changes only at step j, so it is written out
// then; the E rows of X stay in registers).
template <int K>
__device__ __forceinline__ void qr_cols_apply_x(double (&X)[2][K], const double (&u)[K], double tj,
                                                int j, int q, int mr, int rows, int cg, int ng,
                                                int sub, double* Qm) {
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int c = cg + ng * i;
    if (c < q) {
      const double top = c == j ? 1.0 : 0.0;
      double dot = 0.0;
      if (sub == 0) dot = top;
#pragma unroll
      for (int t = 0; t < K; ++t)
        if (sub + 8 * t < mr) dot += u[t] * X[i][t];
      dot = sum8(dot) * tj;
      if (sub == 0) Qm[(size_t)j * rows + c] = top - dot;  // Q1[c][j]
#pragma unroll
      for (int t = 0; t < K; ++t)
        if (sub + 8 * t < mr) X[i][t] -= dot * u[t];
    }
  }
}

template <int K>
__device__ __attribute__((always_inline)) void stacked_qr_cols(double* W, int rows, int q, int mr,
                                                               double* tau, double* Qm) {
  const int tid = threadIdx.x;
  const int ng = (int)blockDim.x >> 3, cg = tid >> 3, sub = tid & 7;
  double X[2][K];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int t = 0; t < K; ++t) X[i][t] = 0.0;
  double E[2][K];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int c = cg + ng * i;
#pragma unroll
    for (int t = 0; t < K; ++t) {
      const int k = sub + 8 * t;
      E[i][t] = (c < q && k < mr) ? W[(size_t)c * rows + q + k] : 0.0;
    }
  }
  __syncthreads();
  if (cg == 0) qr_cols_reflector<K>(E[0], W, rows, q, mr, 0, tau, sub);
  __syncthreads();
  for (int j = 0; j < q - 1; ++j) {
    const double tj = tau[j];
    double u[K];
#pragma unroll
    for (int t = 0; t < K; ++t) {
      const int k = sub + 8 * t;
      u[t] = k < mr ? W[(size_t)j * rows + q + k] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int c = cg + ng * i;
      if (c > j && c < q) {
        double* wc = W + (size_t)c * rows;
        double dot = 0.0;
        if (sub == 0) dot = wc[j];
#pragma unroll
        for (int t = 0; t < K; ++t)
          if (sub + 8 * t < mr) dot += u[t] * E[i][t];
        dot = sum8(dot) * tj;
        if (tj != 0.0) {
          if (sub == 0) wc[j] -= dot;
#pragma unroll
          for (int t = 0; t < K; ++t)
            if (sub + 8 * t < mr) E[i][t] -= dot * u[t];
        }
        if (c == j + 1) qr_cols_reflector<K>(E[i], W, rows, q, mr, j + 1, tau, sub);
      }
    }
    qr_cols_apply_x<K>(X, u, tj, j, q, mr, rows, cg, ng, sub, Qm);
    __syncthreads();
  }
  {
    const int j = q - 1;
    const double tj = tau[j];
    double u[K];
#pragma unroll
    for (int t = 0; t < K; ++t) {
      const int k = sub + 8 * t;
      u[t] = k < mr ? W[(size_t)j * rows + q + k] : 0.0;
    }
    qr_cols_apply_x<K>(X, u, tj, j, q, mr, rows, cg, ng, sub, Qm);
  }
  __syncthreads();
}

// R^-1 column by column (as upper_tri_inverse: for k = q-1 .. 1 every row
// i < k subtracts R_ik (x_k d_k), then rows scale by d_i = 1 / R_ii); steps
// k > c leave column c unchanged.  Needs q <= 2 ng and q <= 8 K.
template <int K>
__device__ __attribute__((always_inline)) void upper_tri_inverse_regs(const double* R, int ld,
                                                                      int q, double* X,
                                                                      double* dinv) {
  const int tid = threadIdx.x;
  const int ng = (int)blockDim.x >> 3, cg = tid >> 3, sub = tid & 7;
  const int base = (tid & (KPC_WAVE - 1)) & ~7;
  for (int i = tid; i < q; i += blockDim.x) dinv[i] = 1.0 / R[(size_t)i * ld + i];
  __syncthreads();
  double x[2][K];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int c = cg + ng * i;
#pragma unroll
    for (int t = 0; t < K; ++t) x[i][t] = (sub + 8 * t == c) ? 1.0 : 0.0;
  }
  for (int k = q - 1; k > 0; --k) {
    const double dk = dinv[k];
    const int tk = k >> 3, sk = k & 7;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int c = cg + ng * i;
      double xk = x[i][0];
#pragma unroll
      for (int t = 1; t < K; ++t)
        if (t == tk) xk = x[i][t];
      xk = __shfl(xk, base + sk, KPC_WAVE);
      if (c < q && k <= c) {
        const double s = xk * dk;
#pragma unroll
        for (int t = 0; t < K; ++t) {
          const int r = sub + 8 * t;
          if (r < k) x[i][t] -= R[(size_t)k * ld + r] * s;
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int c = cg + ng * i;
#pragma unroll
    for (int t = 0; t < K; ++t) {
      const int r = sub + 8 * t;
      if (c < q && r < q) X[(size_t)c * q + r] = r <= c ? x[i][t] * dinv[r] : x[i][t];
    }
  }
  __syncthreads();
}

// Register width for (q, mr): 0 when the LDS forms must be used.  Every
// register form holds 2 K values per lane (two columns per group) and takes
// K <= 6.
__device__ __forceinline__ int regs_width(int q, int mr, int kmax) {
  const int k = max((q + 7) >> 3, (mr + 7) >> 3);
  if (2 * ((int)blockDim.x >> 3) < q) return 0;
  const int w = k <= 2 ? 2 : k <= 4 ? 4 : k <= 6 ? 6 : k <= 8 ? 8 : 0;
  return w <= kmax ? w : 0;
}

// Structured QR; returns true when Q1 (top q rows, ld rows) was formed in Qm
// along the way (register forms), false when the caller must form it.
__device__ __attribute__((always_inline)) bool stacked_qr_fast(double* W, int rows, int q, int mr,
                                                               double* tau, double* Qm) {
  switch (regs_width(q, mr, 6)) {
    case 2: stacked_qr_cols<2>(W, rows, q, mr, tau, Qm); return true;
    case 4: stacked_qr_cols<4>(W, rows, q, mr, tau, Qm); return true;
    case 6: stacked_qr_cols<6>(W, rows, q, mr, tau, Qm); return true;
    default: stacked_qr(W, rows, q, mr, tau); return false;
  }
}

__device__ __attribute__((always_inline)) void stacked_q1_fast(const double* W, int rows, int q,
                                                               int mr, const double* tau,
                                                               double* Qm) {
  switch (regs_width(q, mr, 6)) {
    case 2: stacked_q1_regs<2>(W, rows, q, mr, tau, Qm); break;
    case 4: stacked_q1_regs<4>(W, rows, q, mr, tau, Qm); break;
    case 6: stacked_q1_regs<6>(W, rows, q, mr, tau, Qm); break;
    default: stacked_q1(W, rows, q, mr, tau, Qm);
  }
}

__device__ __attribute__((always_inline)) void upper_tri_inverse_fast(const double* R, int ld, int q,
                                                                      double* X, double* dinv) {
  switch (regs_width(q, 0, 6)) {
    case 2: upper_tri_inverse_regs<2>(R, ld, q, X, dinv); break;
    case 4: upper_tri_inverse_regs<4>(R, ld, q, X, dinv); break;
    case 6: upper_tri_inverse_regs<6>(R, ld, q, X, dinv); break;
    default: upper_tri_inverse(R, ld, q, X, dinv);
  }
}

// ---- one-wave certified-QR evaluation (additive designs: m rank-8 penalty
// blocks, q = 1 + 9 m).  Operation for operation the arithmetic of the
// workgroup path of magic_eval: stacked_qr_cols' per-column dots, updates and
// reflectors (8 lanes per column, lane sub holding E rows sub + 8 t, which
// for rank-8 blocks is penalty block t), Q1 formed alongside, y1 = Q1' y0 per
// column, and the closing block_sum3 of a 256-lane workgroup replayed virtual
// wave by virtual wave (the same partial sums in the same lanes, the same
// shuffle tree, the four wave totals added in the same order).  Only where
// values travel changes: columns c = g + 8 i belong to lane group g, the raw
// tail of column j + 1 is read across lanes from its group once step j has
// updated it and every group forms the reflector itself, so the whole
// factorisation runs without a workgroup barrier and the waves of a
// workgroup can score different trial points at once.

template <int K>
__device__ __forceinline__ void wave_reflector(const double (&ur)[K], double alpha, double (&u)[K],
                                               double& tj, double& beta) {
  double s2 = 0.0;
#pragma unroll
  for (int t = 0; t < K; ++t) s2 += ur[t] * ur[t];
  s2 = sum8(s2);
  double scal = 0.0;
  tj = 0.0;
  beta = alpha;
  if (s2 > 0.0) {
    const double h = fma(alpha, alpha, s2);
    const double r = nrsq(h);
    const double nrm = h * r;
    beta = -copysign(nrm, alpha);
    tj = fma(fabs(alpha), r, 1.0);
    scal = copysign(nrcp(fabs(alpha) + nrm), alpha);
  }
#pragma unroll
  for (int t = 0; t < K; ++t) u[t] = ur[t] * scal;
}

// sum8 of every v[i], i in [I0, I1), stage by stage (the same lane order per
// value as sum8; the stages of different values interleave).
template <int I0, int I1, int N>
__device__ __forceinline__ void sum8_range(double (&v)[N]) {
#pragma unroll
  for (int i = I0; i < I1; ++i) v[i] += dpp_move<0xB1>(v[i]);
#pragma unroll
  for (int i = I0; i < I1; ++i) v[i] += dpp_move<0x4E>(v[i]);
#pragma unroll
  for (int i = I0; i < I1; ++i) v[i] += dpp_move<0x141>(v[i]);
}

// Step j (in segment S = j / 8) of the structured QR on one wave: reflector
// j (u, tj) onto the E columns c > j of slots i >= S (row j of the top block,
// R2[j][c] out) and onto the X columns c <= j of slots i <= S (row j of Q1
// out), then reflector j + 1 from its column's updated tail.  Per column the
// operations and their order are stacked_qr_cols' / qr_cols_apply_x's;
// columns are computed side by side without branches.
template <int K, int NC, int Q, int S>
__device__ __forceinline__ void wave_qr_step(int j, int g, int sub, double (&E)[NC][K],
                                             double (&X)[NC][K], double (&u)[K], double& tj,
                                             const double* Rst, double* R2, int ldr, double* Q1,
                                             int ldq) {
  if constexpr (S < NC) {
    double rjc[NC], dot[NC];
#pragma unroll
    for (int i = S; i < NC; ++i) {
      const int c = g + 8 * i;
      rjc[i] = c < Q ? Rst[c * Q + j] : 0.0;
      dot[i] = 0.0;
      if (sub == 0) dot[i] = rjc[i];
    }
#pragma unroll
    for (int t = 0; t < K; ++t)
#pragma unroll
      for (int i = S; i < NC; ++i) dot[i] += u[t] * E[i][t];
    sum8_range<S, NC>(dot);
    // (columns c <= j and the padding are dead: their E may change freely)
    const bool upd = tj != 0.0;
#pragma unroll
    for (int i = S; i < NC; ++i) {
      const int c = g + 8 * i;
      const double d = dot[i] * tj;
      if (upd) {
#pragma unroll
        for (int t = 0; t < K; ++t) E[i][t] -= d * u[t];
      }
      if (R2 && c > j && c < Q && sub == 0) R2[(size_t)c * ldr + j] = upd ? rjc[i] - d : rjc[i];
    }
    double dx[NC];
#pragma unroll
    for (int i = 0; i <= S; ++i) {
      const int c = g + 8 * i;
      dx[i] = 0.0;
      if (sub == 0) dx[i] = c == j ? 1.0 : 0.0;
    }
#pragma unroll
    for (int t = 0; t < K; ++t)
#pragma unroll
      for (int i = 0; i <= S; ++i) dx[i] += u[t] * X[i][t];
    sum8_range<0, S + 1>(dx);
    // (a column c > j still holds exact +0 and keeps it: d = +0, +0 - (+-0)
    // = +0, and its Q1 entry is the +0 written up front)
#pragma unroll
    for (int i = 0; i <= S; ++i) {
      const int c = g + 8 * i;
      const double d = dx[i] * tj;
      if (c < Q && sub == 0) Q1[(size_t)j * ldq + c] = (c == j ? 1.0 : 0.0) - d;
#pragma unroll
      for (int t = 0; t < K; ++t) X[i][t] -= d * u[t];
    }
    if (j + 1 < Q) {
      const int jn = j + 1, gn = jn & 7;
      const bool next_slot = (jn >> 3) != S;  // column j + 1 opens the next slot
      double ur[K], beta;
#pragma unroll
      for (int t = 0; t < K; ++t) {
        double v = E[S][t];
        if constexpr (S + 1 < NC) v = next_slot ? E[S + 1][t] : v;
        ur[t] = __shfl(v, gn * 8 + sub, KPC_WAVE);
      }
      wave_reflector<K>(ur, Rst[jn * Q + jn], u, tj, beta);
      if (R2 && g == 0 && sub == 0) R2[(size_t)jn * ldr + jn] = beta;
    }
  }
}

// Score of the design at log(sp) = lspv by the calling wave.  Rst (q x q), rSst
// (8m x q) and y0 (Q'y, q) are the fit's staged inputs; nrm the design norms.
// Outputs: R2 (upper triangle, ld ldr; null: not kept), Q1 (ld ldq), y1, and
// res = {score, rss, edf, rank margin, 1} -- or res[4] = 0 when the stacked
// root or the stage-1 rank certificate does not hold (the caller then takes
// the workgroup path, which decides those cases).
template <int K>
__device__ __attribute__((always_inline)) void wave_eval_additive(
    const double* Rst, const double* rSst, const double* y0, const double* nrm,
    const double* lspv, int n, double ynorm2, double rank_tol, double* R2, int ldr, double* Q1,
    int ldq, double* y1, double* res, double* tacc = nullptr) {
  constexpr int Q = 1 + 9 * K;
  constexpr int NC = (Q + 7) / 8;
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const long long tw0 = tacc ? wall_clock64() : 0;
  const int g = lane >> 3, sub = lane & 7;
  double sp[K], sq[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    sp[i] = exp(lspv[i]);
    sq[i] = sqrt(sp[i]);
  }
  double dmax = 0.0;
#pragma unroll
  for (int i = 0; i < K; ++i) dmax = fmax(dmax, sp[i] * nrm[2 + K + i]);
  const double dstop = (double)Q * kEps * dmax;
  bool stacked = true;
#pragma unroll
  for (int i = 0; i < K; ++i) stacked = stacked && sp[i] * nrm[2 + 2 * K + i] > 1e3 * dstop;
  double emax = nrm[0];
#pragma unroll
  for (int i = 0; i < K; ++i) emax += sp[i] * nrm[2 + i];
  const double lb = 1.0 / sqrt(nrm[1] * emax);
  if (!(stacked && isfinite(lb) && lb > 2.0 * rank_tol)) {
    if (lane == 0) res[4] = 0.0;
    return;
  }
  double E[NC][K], X[NC][K];
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    const int c = g + 8 * i;
#pragma unroll
    for (int t = 0; t < K; ++t) {
      E[i][t] = c < Q ? sq[t] * rSst[(sub + 8 * t) * Q + c] : 0.0;
      X[i][t] = 0.0;
    }
  }
  // Q1 is upper triangular (reflector j touches row j of the top block
  // only): row c of Q1' [I; 0] stays exactly +0 before step c, so those
  // entries are written once here and column c of X joins at step c.
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    const int c = g + 8 * i;
    if (c < Q)
      for (int j = sub; j < c; j += 8) Q1[(size_t)j * ldq + c] = 0.0;
  }
  double u[K], tj, beta;
  {
    double ur[K];
#pragma unroll
    for (int t = 0; t < K; ++t) ur[t] = __shfl(E[0][t], sub, KPC_WAVE);
    wave_reflector<K>(ur, Rst[0], u, tj, beta);
    if (R2 && lane == 0) R2[0] = beta;
  }
  const long long tw1 = tacc ? wall_clock64() : 0;
  for (int j = 0; j < Q; ++j) {
    // step j runs in segment j / 8: E columns of slots >= j / 8, X columns of
    // slots <= j / 8 (compile-time ranges; the rest are finished or zero)
    switch (j >> 3) {
      case 0: wave_qr_step<K, NC, Q, 0>(j, g, sub, E, X, u, tj, Rst, R2, ldr, Q1, ldq); break;
      case 1: wave_qr_step<K, NC, Q, 1>(j, g, sub, E, X, u, tj, Rst, R2, ldr, Q1, ldq); break;
      case 2: wave_qr_step<K, NC, Q, 2>(j, g, sub, E, X, u, tj, Rst, R2, ldr, Q1, ldq); break;
      case 3: wave_qr_step<K, NC, Q, 3>(j, g, sub, E, X, u, tj, Rst, R2, ldr, Q1, ldq); break;
      case 4: wave_qr_step<K, NC, Q, 4>(j, g, sub, E, X, u, tj, Rst, R2, ldr, Q1, ldq); break;
      default: wave_qr_step<K, NC, Q, 5>(j, g, sub, E, X, u, tj, Rst, R2, ldr, Q1, ldq); break;
    }
  }
  wave_sync();
  const long long tw2 = tacc ? wall_clock64() : 0;
  // y1 = Q1' y0, one 8-lane group per column
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    const int c = g + 8 * i;
    if (c < Q) {
      double sacc = 0.0;
      for (int r = sub; r < Q; r += 8) sacc += Q1[(size_t)c * ldq + r] * y0[r];
      sacc = sum8(sacc);
      if (sub == 0) y1[c] = sacc;
    }
  }
  wave_sync();
  // y'AAy, edf, y'Ay: the 256-lane partial sums and block_sum3, wave by wave
  double ta = 0.0, tb = 0.0, tc = 0.0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const int vt = 64 * w + lane, cgv = vt >> 3;
    double pa = 0.0, pe = 0.0, py = 0.0;
    for (int r = cgv; r < Q; r += 32) {
      double v = 0.0;
      for (int c = sub; c < Q; c += 8) {
        const double uu = Q1[(size_t)c * ldq + r];
        v += uu * y1[c];
        pe += uu * uu;
      }
      v = sum8(v);
      if (sub == 0) pa += v * v;
    }
    if (vt < Q) py += y1[vt] * y1[vt];
    pa = wave_sum(pa);
    pe = wave_sum(pe);
    py = wave_sum(py);
    ta += pa;
    tb += pe;
    tc += py;
  }
  if (lane == 0) {
    const double yAAy = ta, edf = tb, yAy = tc;
    double rss = ynorm2 - 2.0 * yAy + yAAy;
    if (rss < 0.0) rss = 0.0;
    const double delta = (double)n - edf;
    res[0] = delta > 1e-8 ? (double)n * rss / (delta * delta) : __builtin_inf();
    res[1] = rss;
    res[2] = edf;
    res[3] = lb / rank_tol - 1.0;
    res[4] = 1.0;
    if (tacc) {  // phase ticks: set-up, QR with Q1, y1 and the sums
      const long long tw3 = wall_clock64();
      tacc[0] += (double)(tw1 - tw0);
      tacc[1] += (double)(tw2 - tw1);
      tacc[2] += (double)(tw3 - tw2);
    }
  }
  wave_sync();
}

// Two trial points on one wave: wave_qr_step for points a and b side by
// side (the design's R entries are loaded once for both).  Each point's
// operations and their order are exactly wave_qr_step's, so each point's
// outputs equal a one-point evaluation's; the two dependency chains give the
// wave twice the independent work per step.  R2 is written for point a only.
template <int K, int NC, int Q, int S>
__device__ __forceinline__ void wave_qr_step2(int j, int g, int sub, double (&E)[2][NC][K],
                                              double (&X)[2][NC][K], double (&u)[2][K],
                                              double (&tj)[2], const double* Rst, double* R2,
                                              int ldr, double* Q1a, int lda, double* Q1b,
                                              int ldb) {
  if constexpr (S < NC) {
    double rjc[NC], dot[2][NC];
#pragma unroll
    for (int i = S; i < NC; ++i) {
      const int c = g + 8 * i;
      rjc[i] = c < Q ? Rst[c * Q + j] : 0.0;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        dot[p][i] = 0.0;
        if (sub == 0) dot[p][i] = rjc[i];
      }
    }
#pragma unroll
    for (int t = 0; t < K; ++t)
#pragma unroll
      for (int i = S; i < NC; ++i)
#pragma unroll
        for (int p = 0; p < 2; ++p) dot[p][i] += u[p][t] * E[p][i][t];
    sum8_range<S, NC>(dot[0]);
    sum8_range<S, NC>(dot[1]);
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const bool upd = tj[p] != 0.0;
#pragma unroll
      for (int i = S; i < NC; ++i) {
        const int c = g + 8 * i;
        const double d = dot[p][i] * tj[p];
        if (upd) {
#pragma unroll
          for (int t = 0; t < K; ++t) E[p][i][t] -= d * u[p][t];
        }
        if (p == 0 && R2 && c > j && c < Q && sub == 0)
          R2[(size_t)c * ldr + j] = upd ? rjc[i] - d : rjc[i];
      }
    }
    double dx[2][NC];
#pragma unroll
    for (int i = 0; i <= S; ++i) {
      const int c = g + 8 * i;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        dx[p][i] = 0.0;
        if (sub == 0) dx[p][i] = c == j ? 1.0 : 0.0;
      }
    }
#pragma unroll
    for (int t = 0; t < K; ++t)
#pragma unroll
      for (int i = 0; i <= S; ++i)
#pragma unroll
        for (int p = 0; p < 2; ++p) dx[p][i] += u[p][t] * X[p][i][t];
    sum8_range<0, S + 1>(dx[0]);
    sum8_range<0, S + 1>(dx[1]);
#pragma unroll
    for (int i = 0; i <= S; ++i) {
      const int c = g + 8 * i;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const double d = dx[p][i] * tj[p];
        if (c < Q && sub == 0) {
          if (p == 0)
            Q1a[(size_t)j * lda + c] = (c == j ? 1.0 : 0.0) - d;
          else
            Q1b[(size_t)j * ldb + c] = (c == j ? 1.0 : 0.0) - d;
        }
#pragma unroll
        for (int t = 0; t < K; ++t) X[p][i][t] -= d * u[p][t];
      }
    }
    if (j + 1 < Q) {
      const int jn = j + 1, gn = jn & 7;
      const bool next_slot = (jn >> 3) != S;  // column j + 1 opens the next slot
      const double rdiag = Rst[jn * Q + jn];
      double ur[2][K], beta[2];
#pragma unroll
      for (int t = 0; t < K; ++t)
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          double v = E[p][S][t];
          if constexpr (S + 1 < NC) v = next_slot ? E[p][S + 1][t] : v;
          ur[p][t] = __shfl(v, gn * 8 + sub, KPC_WAVE);
        }
      wave_reflector<K>(ur[0], rdiag, u[0], tj[0], beta[0]);
      wave_reflector<K>(ur[1], rdiag, u[1], tj[1], beta[1]);
      if (R2 && g == 0 && sub == 0) R2[(size_t)jn * ldr + jn] = beta[0];
    }
  }
}

// wave_eval_additive for two points at once (lspa, lspb), each with its own
// Q1 / y1 / res outputs (R2 and the phase ticks for point a only); each
// point's outputs are bit-identical to a one-point evaluation of it.
template <int K>
__device__ __attribute__((always_inline)) void wave_eval_additive2(
    const double* Rst, const double* rSst, const double* y0, const double* nrm,
    const double* lspa, const double* lspb, int n, double ynorm2, double rank_tol, double* R2,
    int ldr, double* Q1a, int lda, double* y1a, double* resa, double* tacc, double* Q1b, int ldb,
    double* y1b, double* resb) {
  constexpr int Q = 1 + 9 * K;
  constexpr int NC = (Q + 7) / 8;
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const long long tw0 = tacc ? wall_clock64() : 0;
  const int g = lane >> 3, sub = lane & 7;
  const double* lspv[2] = {lspa, lspb};
  double* res[2] = {resa, resb};
  double lb[2];
  bool ok[2];
  double E[2][NC][K], X[2][NC][K];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    double sp[K], sq[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
      sp[i] = exp(lspv[p][i]);
      sq[i] = sqrt(sp[i]);
    }
    double dmax = 0.0;
#pragma unroll
    for (int i = 0; i < K; ++i) dmax = fmax(dmax, sp[i] * nrm[2 + K + i]);
    const double dstop = (double)Q * kEps * dmax;
    bool stacked = true;
#pragma unroll
    for (int i = 0; i < K; ++i) stacked = stacked && sp[i] * nrm[2 + 2 * K + i] > 1e3 * dstop;
    double emax = nrm[0];
#pragma unroll
    for (int i = 0; i < K; ++i) emax += sp[i] * nrm[2 + i];
    lb[p] = 1.0 / sqrt(nrm[1] * emax);
    ok[p] = stacked && isfinite(lb[p]) && lb[p] > 2.0 * rank_tol;
    if (!ok[p] && lane == 0) res[p][4] = 0.0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      const int c = g + 8 * i;
#pragma unroll
      for (int t = 0; t < K; ++t) {
        E[p][i][t] = c < Q ? sq[t] * rSst[(sub + 8 * t) * Q + c] : 0.0;
        X[p][i][t] = 0.0;
      }
    }
  }
  if (!ok[0] && !ok[1]) return;
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    const int c = g + 8 * i;
    if (c < Q)
      for (int j = sub; j < c; j += 8) {
        Q1a[(size_t)j * lda + c] = 0.0;
        Q1b[(size_t)j * ldb + c] = 0.0;
      }
  }
  double u[2][K], tj[2];
  {
    double ur[2][K], beta[2];
#pragma unroll
    for (int t = 0; t < K; ++t)
#pragma unroll
      for (int p = 0; p < 2; ++p) ur[p][t] = __shfl(E[p][0][t], sub, KPC_WAVE);
    wave_reflector<K>(ur[0], Rst[0], u[0], tj[0], beta[0]);
    wave_reflector<K>(ur[1], Rst[0], u[1], tj[1], beta[1]);
    if (R2 && lane == 0) R2[0] = beta[0];
  }
  const long long tw1 = tacc ? wall_clock64() : 0;
  for (int j = 0; j < Q; ++j) {
    switch (j >> 3) {
      case 0: wave_qr_step2<K, NC, Q, 0>(j, g, sub, E, X, u, tj, Rst, R2, ldr, Q1a, lda, Q1b, ldb); break;
      case 1: wave_qr_step2<K, NC, Q, 1>(j, g, sub, E, X, u, tj, Rst, R2, ldr, Q1a, lda, Q1b, ldb); break;
      case 2: wave_qr_step2<K, NC, Q, 2>(j, g, sub, E, X, u, tj, Rst, R2, ldr, Q1a, lda, Q1b, ldb); break;
      case 3: wave_qr_step2<K, NC, Q, 3>(j, g, sub, E, X, u, tj, Rst, R2, ldr, Q1a, lda, Q1b, ldb); break;
      case 4: wave_qr_step2<K, NC, Q, 4>(j, g, sub, E, X, u, tj, Rst, R2, ldr, Q1a, lda, Q1b, ldb); break;
      default: wave_qr_step2<K, NC, Q, 5>(j, g, sub, E, X, u, tj, Rst, R2, ldr, Q1a, lda, Q1b, ldb); break;
    }
  }
  wave_sync();
  const long long tw2 = tacc ? wall_clock64() : 0;
  double* Q1v[2] = {Q1a, Q1b};
  const int ldv[2] = {lda, ldb};
  double* y1v[2] = {y1a, y1b};
  // y1 = Q1' y0, one 8-lane group per column
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    const int c = g + 8 * i;
    if (c < Q) {
      double sacc[2] = {0.0, 0.0};
      for (int r = sub; r < Q; r += 8)
#pragma unroll
        for (int p = 0; p < 2; ++p) sacc[p] += Q1v[p][(size_t)c * ldv[p] + r] * y0[r];
      sacc[0] = sum8(sacc[0]);
      sacc[1] = sum8(sacc[1]);
      if (sub == 0) {
        y1a[c] = sacc[0];
        y1b[c] = sacc[1];
      }
    }
  }
  wave_sync();
  // y'AAy, edf, y'Ay: the 256-lane partial sums and block_sum3, wave by wave
  double ta[2] = {0.0, 0.0}, tb[2] = {0.0, 0.0}, tc[2] = {0.0, 0.0};
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const int vt = 64 * w + lane, cgv = vt >> 3;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      double pa = 0.0, pe = 0.0, py = 0.0;
      for (int r = cgv; r < Q; r += 32) {
        double v = 0.0;
        for (int c = sub; c < Q; c += 8) {
          const double uu = Q1v[p][(size_t)c * ldv[p] + r];
          v += uu * y1v[p][c];
          pe += uu * uu;
        }
        v = sum8(v);
        if (sub == 0) pa += v * v;
      }
      if (vt < Q) py += y1v[p][vt] * y1v[p][vt];
      pa = wave_sum(pa);
      pe = wave_sum(pe);
      py = wave_sum(py);
      ta[p] += pa;
      tb[p] += pe;
      tc[p] += py;
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      if (!ok[p]) continue;
      const double yAAy = ta[p], edf = tb[p], yAy = tc[p];
      double rss = ynorm2 - 2.0 * yAy + yAAy;
      if (rss < 0.0) rss = 0.0;
      const double delta = (double)n - edf;
      res[p][0] = delta > 1e-8 ? (double)n * rss / (delta * delta) : __builtin_inf();
      res[p][1] = rss;
      res[p][2] = edf;
      res[p][3] = lb[p] / rank_tol - 1.0;
      res[p][4] = 1.0;
    }
    if (tacc) {  // phase ticks: set-up, QR with Q1, y1 and the sums
      const long long tw3 = wall_clock64();
      tacc[0] += (double)(tw1 - tw0);
      tacc[1] += (double)(tw2 - tw1);
      tacc[2] += (double)(tw3 - tw2);
    }
  }
  wave_sync();
}

// Total penalty sum_i sp_i S_i = sum_i sp_i rS_i rS_i' (q x q, column-major).
__device__ __attribute__((always_inline)) void aggregate_penalty(const double* D, const DesignLayout& DL, int q, int m, int rk,
                                  const double* spv, double* agg) {
  for (int e = threadIdx.x; e < q * q; e += blockDim.x) {
    const int r = e % q, c = e / q;
    double a = 0.0;
    for (int i = 0; i < m; ++i) {
      const double* R0 = D + DL.rS + (size_t)i * rk * q;
      double v = 0.0;
      for (int k2 = 0; k2 < rk; ++k2) v += R0[(size_t)k2 * q + r] * R0[(size_t)k2 * q + c];
      a += v * spv[i];
    }
    agg[e] = a;
  }
  __syncthreads();
}

// Copy [R; root] into the augmented buffer (rows = q + rank).
__device__ __attribute__((always_inline)) void build_augmented(const MagicLds& L, const double* D, const DesignLayout& DL, int q,
                                int rank, const double* root) {
  const int tid = threadIdx.x;
  const int rows = q + rank;
  for (int e = tid; e < rows * q; e += blockDim.x) {
    const int r = e % rows, c = e / rows;
    L.aug[e] = r < q ? D[DL.R + (size_t)c * q + r] : 0.0;
  }
  __syncthreads();
  for (int e = tid; e < rank * q; e += blockDim.x) {
    const int r = e % rank, col = e / rank;
    if (r <= col) L.aug[(size_t)L.piv[col] * rows + q + r] = root[(size_t)col * q + r];
  }
  __syncthreads();
}

// GCV core at log(sp) = lspv.  Results: dctl[0]=rss dctl[1]=edf dctl[2]=score
// dctl[6]=rank-decision margin, ictl[0]=nr ictl[1]=rank (aggregate root)
// ictl[2]=status ictl[13]=basis (1: QR, 0: SVD), y1.
//
// mgcv magic takes the SVD [R; E] = U D V' and keeps the singular values above
// d_max sqrt(eps) (the numerical rank nr).  When the rank is certifiably full
// -- d_min / d_max >= 1 / (|R2|_F |R2^-1|_F) > 2 rank_tol for the triangular
// factor R2 of [R; E] -- every quantity magic derives from U1 (first q rows of
// U) and V D^-1 is invariant under the orthogonal change of basis U1 = Q1 W,
// V D^-1 = R2^-1 W, so the structured QR (q reflectors of length 1 + rank)
// replaces the Jacobi SVD; otherwise the SVD decides the rank exactly.
__device__ __attribute__((always_inline)) void magic_eval(const MagicLds& L, const MagicScratch& S, const double* D,
                           const DesignLayout& DL, int n, int q, int m, int rk,
                           const double* lspv, double ynorm2, double rank_tol) {
  const int tid = threadIdx.x;
  const int lane = tid & (KPC_WAVE - 1);
  const int wid = tid / KPC_WAVE;
  const int nw = blockDim.x / KPC_WAVE;
  const double* spv = L.spv;
  const long long te0 = wall_clock64();
  if (tid < m) {
    const double e = exp(lspv[tid]);
    L.spv[tid] = e;
    L.spv[m + tid] = sqrt(e);
  }
  __syncthreads();
  // The total-penalty root of mgcv (pivoted Cholesky, default tolerance
  // q eps max diag) drops every pivot <= dstop.  The penalties act on disjoint
  // coefficient blocks, so the pivots of block i are sp_i times the pivots of
  // chol(S_i); when even the smallest of them clears dstop by a wide margin
  // the root has full rank m rk and [R; sqrt(sp_i) rS_i'] has the same Gram
  // matrix as [R; root], i.e. the same singular values and U1, V, D.
  const int mr = m * rk;
  double dmax = 0.0;
  for (int i = 0; i < m; ++i) dmax = fmax(dmax, spv[i] * L.nrm[2 + m + i]);
  const double dstop = (double)q * kEps * dmax;
  bool stacked = true;
  for (int i = 0; i < m; ++i) stacked = stacked && spv[i] * L.nrm[2 + 2 * m + i] > 1e3 * dstop;
  long long tp0 = wall_clock64(), tp1 = tp0, tp2 = tp0;
  int rows, erows = mr;
  if (tid == 0) L.ictl[15] = 0;
  bool qr_ok = false;
  double lb = 0.0;
  if (stacked) {
    // full-rank certificate, stage 1 from the design norms: sigma_min([R; E])
    // >= sigma_min(R) >= 1 / |R^-1|_F, sigma_max^2 <= |R|_F^2 + sum_i sp_i |rS_i|_F^2
    double emax = L.nrm[0];
    for (int i = 0; i < m; ++i) emax += spv[i] * L.nrm[2 + i];
    lb = 1.0 / sqrt(L.nrm[1] * emax);
    qr_ok = isfinite(lb) && lb > 2.0 * rank_tol;
    rows = q + mr;
    for (int e = tid; e < rows * q; e += blockDim.x) {
      const int r = e % rows, c = e / rows;
      L.aug[e] = r < q ? D[DL.R + (size_t)c * q + r]
                       : spv[m + (r - q) / rk] * D[DL.rS + (size_t)(r - q) * q + c];
    }
    __syncthreads();
  } else {
    // root with dropped pivots: exactly mgcv's [R; root]
    double* agg = L.X4;
    aggregate_penalty(D, DL, q, m, rk, spv, agg);
    __syncthreads();
    if (wid == 0) {
      const int r = pstf2_upper_wave(agg, q, L.piv, L.wk);
      if (lane == 0) L.ictl[14] = r;
    }
    __syncthreads();
    erows = L.ictl[14];
    rows = q + erows;
    build_augmented(L, D, DL, q, erows, agg);
  }
  const long long tq0 = wall_clock64();
  const bool q1_done = stacked_qr_fast(L.aug, rows, q, erows, L.wk, L.Qm);
  const long long tq1 = wall_clock64();
  if (!qr_ok) {
    // stage 2 from the factor itself: 1 / (|R2|_F |R2^-1|_F)
    upper_tri_inverse_fast(L.aug, rows, q, L.V, L.wk + q);
    double pr = 0.0, pi = 0.0;
    for (int e = tid; e < q * q; e += blockDim.x) {
      const int r = e % q, c = e / q;
      if (r <= c) {
        const double a = L.aug[(size_t)c * rows + r], b = L.V[e];
        pr += a * a;
        pi += b * b;
      }
    }
    const double nr2 = block_sum(pr, L.red);
    const double ni2 = block_sum(pi, L.red);
    lb = 1.0 / sqrt(nr2 * ni2);
    qr_ok = isfinite(lb) && lb > 2.0 * rank_tol;
    if (qr_ok && tid == 0) L.ictl[15] = 1;  // R2 inverted already
  }
  const long long tq2 = wall_clock64();
  if (qr_ok && !q1_done) stacked_q1_fast(L.aug, rows, q, erows, L.wk, L.Qm);
  const long long tq3 = wall_clock64();
  if (tid == 0) {
    L.dctl[8] += (double)(tq0 - te0);   // penalty scaling + augmented build
    L.dctl[9] += (double)(tq1 - tq0);   // stacked QR
    L.dctl[10] += (double)(tq2 - tq1);  // rank certificate from the factor
    L.dctl[11] += (double)(tq3 - tq2);  // Q1
  }
  if (qr_ok) {
    // ---- QR path on [R; sqrt(sp_i) rS_i'] (same Gram matrix as [R; E])
    if (tid == 0) {
      L.ictl[0] = q;
      L.ictl[1] = erows;
      L.ictl[13] = 1;
      L.dctl[6] = lb / rank_tol - 1.0;
    }
    // y1 = Q1' y0, one aligned 8-lane group per column
    {
      const int ng = (int)blockDim.x >> 3, cg = tid >> 3, sub = tid & 7;
      for (int c = cg; c < q; c += ng) {
        const double* u = L.Qm + (size_t)c * rows;
        double sacc = 0.0;
        for (int r = sub; r < q; r += 8) sacc += u[r] * L.y[r];
        sacc = sum8(sacc);
        if (sub == 0) L.y1[c] = sacc;
      }
    }
    __syncthreads();
  } else {
    // ---- exact path: Jacobi SVD of the triangular factor R2 of the matrix
    // just factored ([R; E] or [R; root], the same Gram matrix).  [R; E] =
    // Q R2 and R2 = U_R D V' give mgcv's D and V directly and U1 = Q1 U_R
    // (first q rows of U), so the one-sided Jacobi runs on q rows instead of
    // q + rank and the root is not rebuilt.
    const int rank = erows;
    tp1 = wall_clock64();
    if (!q1_done) stacked_q1_fast(L.aug, rows, q, erows, L.wk, L.Qm);
    for (int e = tid; e < q * q; e += blockDim.x) {
      const int r = e % q, c = e / q;
      L.X4[e] = r <= c ? L.aug[(size_t)c * rows + r] : 0.0;
    }
    __syncthreads();
    tp2 = wall_clock64();
    jacobi_svd(L.X4, q, q, L.V, L.d, L.ord, L.ictl);
    for (int e = tid; e < q * q; e += blockDim.x) {
      const int r = e % q, c = e / q;
      const double* ur = L.X4 + (size_t)c * q;
      double v = 0.0;
      for (int k = 0; k < q; ++k) v += L.Qm[(size_t)k * rows + r] * ur[k];
      L.aug[(size_t)c * rows + r] = v;
    }
    __syncthreads();
    const double d0 = L.d[L.ord[0]];
    const double thr = d0 * rank_tol;
    int cnt = 0;
    double rm = 1e300;
    for (int i = tid; i < q; i += blockDim.x) {
      const double di = L.d[i];
      cnt += di >= thr;
      rm = fmin(rm, fabs(di / d0 - rank_tol) / rank_tol);
    }
    const double nrs = block_sum((double)cnt, L.red);
    const double rmin = -block_max(-rm, L.red);
    if (tid == 0) {
      L.ictl[0] = (int)nrs;
      L.ictl[1] = rank;
      L.ictl[13] = 0;
      L.dctl[6] = rmin;
    }
    __syncthreads();
    const int nr = L.ictl[0];
    // y1 = U1' y0
    for (int c = wid; c < nr; c += nw) {
      const double* u = L.aug + (size_t)L.ord[c] * rows;
      double sacc = 0.0;
      for (int r = lane; r < q; r += KPC_WAVE) sacc += u[r] * L.y[r];
      sacc = wave_sum(sacc);
      if (lane == 0) L.y1[c] = sacc;
    }
    __syncthreads();
  }
  long long tp3 = wall_clock64();
  if (tid == 0) {
    L.dctl[26] += (double)(tp1 - tp0);
    L.dctl[27] += (double)(tp3 - tp2);
  }
  const int nr = L.ictl[0];
  const bool qr = L.ictl[13] != 0;
  rows = q + L.ictl[1];
  double p_yaay = 0.0, p_edf = 0.0, p_yay = 0.0;
  if (qr) {
    // (Q1 y1)_r and the squares of row r of Q1, one aligned 8-lane group per row
    const int ng = (int)blockDim.x >> 3, cg = tid >> 3, sub = tid & 7;
    for (int r = cg; r < q; r += ng) {
      double v = 0.0;
      for (int c = sub; c < nr; c += 8) {
        const double u = L.Qm[(size_t)c * rows + r];
        v += u * L.y1[c];
        p_edf += u * u;
      }
      v = sum8(v);
      if (sub == 0) p_yaay += v * v;
    }
  } else {
    for (int r = tid; r < q; r += blockDim.x) {
      double v = 0.0;
      for (int c = 0; c < nr; ++c) {
        const double u = L.aug[(size_t)L.ord[c] * rows + r];
        v += u * L.y1[c];
        p_edf += u * u;
      }
      p_yaay += v * v;
    }
  }
  for (int c = tid; c < nr; c += blockDim.x) p_yay += L.y1[c] * L.y1[c];
  block_sum3(p_yaay, p_edf, p_yay, L.red);
  const double yAAy = p_yaay, edf = p_edf, yAy = p_yay;
  if (tid == 0) {
    double rss = ynorm2 - 2.0 * yAy + yAAy;
    if (rss < 0.0) rss = 0.0;
    const double delta = (double)n - edf;
    L.dctl[0] = rss;
    L.dctl[1] = edf;
    L.dctl[2] = delta > 1e-8 ? (double)n * rss / (delta * delta) : __builtin_inf();
    L.ictl[2] = 0;
    L.dctl[20] += (double)(wall_clock64() - te0);
  }
  __syncthreads();
}

// Gradient and Hessian of the GCV score at the point last evaluated.
__device__ __attribute__((always_inline)) void magic_derivs(const MagicLds& L, const MagicScratch& S, const double* D,
                             const DesignLayout& DL, int n, int q, int m, int rk,
                             const double* lspv) {
  const int tid = threadIdx.x;
  const int nr = L.ictl[0];
  const int rows = q + L.ictl[1];
  const int mr = m * rk;
  const long long td0 = wall_clock64();
  const bool qr = L.ictl[13] != 0;
  // G = U1' U1 (nr x nr) first: after it the augmented buffer is free and
  // holds T and GT, so the whole derivative working set stays in LDS.  On the
  // QR basis U1 -> Q1 and D^-1 V1' -> R2^-T (same traces and products).  The
  // dense products run on FP64 matrix cores (wg_gemm_tn).
  double* G = L.X4;
  double* T = L.aug;                          // nr x mr (mr <= q)
  double* GT = L.aug + (size_t)q * q;         // nr x mr
  if (qr) {
    wg_gemm_tn(L.Qm, rows, L.Qm, rows, G, nr, nr, nr, q);
    __syncthreads();
    if (!L.ictl[15]) upper_tri_inverse_fast(L.aug, rows, q, L.V, L.wk);  // before T overwrites R2
    // T = R2^-T rS (R2^-1 is upper triangular with exact zeros below)
    wg_gemm_tn(L.V, q, D + DL.rS, q, T, nr, nr, mr, q);
  } else {
    for (int e = tid; e < nr * nr; e += blockDim.x) {
      const int i = e % nr, j = e / nr;
      const double* ui = L.aug + (size_t)L.ord[i] * rows;
      const double* uj = L.aug + (size_t)L.ord[j] * rows;
      double s = 0.0;
      for (int r = 0; r < q; ++r) s += ui[r] * uj[r];
      G[e] = s;
    }
    __syncthreads();
    // T[c][a] = (V1' rS)[c][a] / d_c
    for (int e = tid; e < nr * mr; e += blockDim.x) {
      const int c = e % nr, a = e / nr;
      const double* rs = D + DL.rS + (size_t)a * q;
      const double* v = L.V + (size_t)L.ord[c] * q;
      double s = 0.0;
      for (int k = 0; k < q; ++k) s += v[k] * rs[k];
      T[(size_t)a * nr + c] = s / L.d[L.ord[c]];
    }
  }
  __syncthreads();
  // GT = G T (G symmetric), gy = G y1
  wg_gemm_tn(G, nr, T, nr, GT, nr, nr, mr, nr);
  // short reductions below: one aligned 8-lane group per output (DPP sums)
  const int ngr = (int)blockDim.x >> 3, cgr = tid >> 3, sub8 = tid & 7;
  for (int c = cgr; c < nr; c += ngr) {
    double s = 0.0;
    for (int k = sub8; k < nr; k += 8) s += G[(size_t)k * nr + c] * L.y1[k];
    s = sum8(s);
    if (sub8 == 0) L.gyl[c] = s;
  }
  __syncthreads();
  // P = T'T and Qx = T' G T (mr x mr) in the Q1 area (free from here on):
  // tr(M_j K_i) = sum_{a in j, b in i} P[a][b] Qx[b][a], tr(K_i) = sum_{a in i} Qx[a][a]
  double* P = L.Qm;
  double* Qx = L.Qm + (size_t)mr * mr;
  wg_gemm_tn(T, nr, T, nr, P, mr, mr, mr, nr);
  wg_gemm_tn(T, nr, GT, nr, Qx, mr, mr, mr, nr);
  // per penalty: My = T_i (T_i' y1), Ky = T_i (T_i' G y1), yK = G My; the
  // column projections T_a' y1 and T_a' G y1 are formed once
  for (int a = cgr; a < mr; a += ngr) {
    const double* Ta = T + (size_t)a * nr;
    double ty = 0.0, tg = 0.0;
    for (int c = sub8; c < nr; c += 8) { ty += Ta[c] * L.y1[c]; tg += Ta[c] * L.gyl[c]; }
    ty = sum8(ty);
    tg = sum8(tg);
    if (sub8 == 0) {
      L.projl[2 * a] = ty;
      L.projl[2 * a + 1] = tg;
    }
  }
  __syncthreads();
  for (int e = tid; e < m * nr; e += blockDim.x) {
    const int i = e / nr, o = e % nr;
    double my = 0.0, ky = 0.0;
    for (int a = 0; a < rk; ++a) {
      const double t = T[(size_t)(i * rk + a) * nr + o];
      my += t * L.projl[2 * (i * rk + a)];
      ky += t * L.projl[2 * (i * rk + a) + 1];
    }
    L.My[(size_t)i * nr + o] = my;
    L.Ky[(size_t)i * nr + o] = ky;
  }
  __syncthreads();
  for (int e = cgr; e < m * nr; e += ngr) {
    const int i = e / nr, o = e % nr;
    double s = 0.0;
    for (int k = sub8; k < nr; k += 8) s += G[(size_t)k * nr + o] * L.My[(size_t)i * nr + k];
    s = sum8(s);
    if (sub8 == 0) L.yK[(size_t)i * nr + o] = s;
  }
  const long long td1 = wall_clock64();
  // traces: one wave per (i, j <= i) pair and per tr(K_i), no workgroup barrier
  {
    const int lane = tid & (KPC_WAVE - 1);
    const int wid = tid / KPC_WAVE;
    const int nw = blockDim.x / KPC_WAVE;
    const int npair = m * (m + 1) / 2;
    for (int e = wid; e < npair + m; e += nw) {
      double part = 0.0;
      int i = 0, j = 0;
      if (e < npair) {
        int rem = e;
        while (rem > i) {
          rem -= i + 1;
          ++i;
        }
        j = rem;
        for (int u = lane; u < rk * rk; u += KPC_WAVE) {
          const int ja = j * rk + u % rk, ib = i * rk + u / rk;
          part += P[(size_t)ib * mr + ja] * Qx[(size_t)ja * mr + ib];
        }
      } else {
        i = e - npair;
        for (int a = lane; a < rk; a += KPC_WAVE) {
          const int ia = i * rk + a;
          part += Qx[(size_t)ia * mr + ia];
        }
      }
      part = wave_sum(part);
      if (lane == 0) {
        if (e < npair)
          L.tr[j * m + i] = part;
        else
          L.tr[m * m + i] = part;
      }
    }
  }
  __syncthreads();
  const long long td2 = wall_clock64();
  if (tid == 0) {
    L.dctl[12] += (double)(td1 - td0);  // G, T, G T, P, Qx, projections
    L.dctl[13] += (double)(td2 - td1);  // traces
  }
  // d|y-Ay|^2 terms: one 8-lane group per i (dnorm) and per pair j <= i (d2n)
  const int npairs = m * (m + 1) / 2;
  for (int e = cgr; e < m + npairs; e += ngr) {
    if (e < m) {
      const int i = e;
      double nd = 0.0;
      for (int c = sub8; c < nr; c += 8)
        nd += L.y1[c] * (L.My[(size_t)i * nr + c] - L.Ky[(size_t)i * nr + c]);
      nd = sum8(nd);
      if (sub8 == 0) L.dn[i] = 2.0 * exp(lspv[i]) * nd;
    } else {
      int i = 0, rem = e - m;
      while (rem > i) {
        rem -= i + 1;
        ++i;
      }
      const int j = rem;
      double nv = 0.0;
      for (int c = sub8; c < nr; c += 8) {
        const double myi = L.My[(size_t)i * nr + c], myj = L.My[(size_t)j * nr + c];
        nv += myi * L.Ky[(size_t)j * nr + c] + myj * L.Ky[(size_t)i * nr + c] - 2.0 * myi * myj +
              L.yK[(size_t)i * nr + c] * myj;
      }
      nv = sum8(nv);
      nv *= 2.0 * exp(lspv[i] + lspv[j]);
      if (sub8 == 0) {
        L.d2n[j * m + i] = nv;
        L.d2n[i * m + j] = nv;
      }
    }
  }
  __syncthreads();
  if (tid == 0) {
    double* dnorm = L.dn;
    double* d2n = L.d2n;
    double* ddelta = L.syw;   // m (free until the eigen-decomposition)
    double* d2d = L.syA;      // m x m (likewise)
    for (int i = 0; i < m; ++i) {
      const double li = exp(lspv[i]);
      ddelta[i] = li * L.tr[m * m + i];
      for (int j = 0; j <= i; ++j) {
        const double dv = -2.0 * exp(lspv[i] + lspv[j]) * L.tr[j * m + i];
        d2d[j * m + i] = dv;
        d2d[i * m + j] = dv;
      }
      d2d[i * m + i] += ddelta[i];
      d2n[i * m + i] += dnorm[i];
    }
    const double rss = L.dctl[0], edf = L.dctl[1];
    const double delta = (double)n - edf;
    const double ss = (double)n / (delta * delta);
    const double ds = 2.0 * ss * rss / delta;
    const double cs = -2.0 * ss / delta;
    const double cv = 3.0 * ds / delta;
    for (int i = 0; i < m; ++i) L.cgrad[i] = ss * dnorm[i] - ds * ddelta[i];
    L.dctl[29] += (double)(wall_clock64() - td0);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j <= i; ++j) {
        const double v = cs * (ddelta[j] * dnorm[i] + ddelta[i] * dnorm[j]) + ss * d2n[j * m + i] +
                         cv * ddelta[i] * ddelta[j] - ds * d2d[j * m + i];
        L.hess[j * m + i] = v;
        L.hess[i * m + j] = v;
      }
    L.dctl[21] += (double)(wall_clock64() - td0);
  }
  __syncthreads();
}

// Symmetric eigen of the small Hessian by cyclic Jacobi (the oracle's rotation
// order and arithmetic), run by one wave: each rotation's element updates are
// spread over the lanes, the rotation itself is computed redundantly by every
// lane, and the off/total test is summed by lane 0 in the serial order.  A
// sweep is a deterministic function of (A, Z), so once a sweep changes no bit
// every later sweep is the identity and the loop stops with exactly the state
// the full sweep budget reaches.
__device__ __forceinline__ bool bits_differ(double x, double y) {
  return __double_as_longlong(x) != __double_as_longlong(y);
}

// small_syev_wave for a compile-time M <= 4 with A and Z in registers: every
// lane of the wave runs the same rotations with the same operations in the
// same order (each rotation's column and row updates, the stop test's sums,
// the bitwise fixed point), so w and Z are bit-identical to the LDS form
// without its per-rotation LDS round trips and wave syncs.
template <int M>
__device__ __attribute__((always_inline)) void small_syev_regs(const double* Hin, double* w,
                                                               double* Z) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  double A[M * M], V[M * M];
#pragma unroll
  for (int i = 0; i < M * M; ++i) {
    A[i] = Hin[i];
    V[i] = (i % M == i / M) ? 1.0 : 0.0;
  }
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
#pragma unroll
    for (int i = 0; i < M * M; ++i) {
      tot += A[i] * A[i];
      if (i % M != i / M) off += A[i] * A[i];
    }
    if (off <= 1e-300 || off <= kEps * kEps * tot * 1e-4) break;
    bool changed = false;
#pragma unroll
    for (int p = 0; p < M - 1; ++p)
#pragma unroll
      for (int q2 = p + 1; q2 < M; ++q2) {
        const double apq = A[q2 * M + p];
        if (apq == 0.0) continue;
        const double app = A[p * M + p], aqq = A[q2 * M + q2];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = copysign(1.0, theta) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
        for (int k = 0; k < M; ++k) {
          const double akp = A[p * M + k], akq = A[q2 * M + k];
          const double np = c * akp - s * akq, nq = s * akp + c * akq;
          changed |= bits_differ(np, akp) || bits_differ(nq, akq);
          A[p * M + k] = np;
          A[q2 * M + k] = nq;
        }
#pragma unroll
        for (int k = 0; k < M; ++k) {
          const double apk = A[k * M + p], aqk = A[k * M + q2];
          const double np = c * apk - s * aqk, nq = s * apk + c * aqk;
          changed |= bits_differ(np, apk) || bits_differ(nq, aqk);
          A[k * M + p] = np;
          A[k * M + q2] = nq;
        }
#pragma unroll
        for (int k = 0; k < M; ++k) {
          const double zkp = V[p * M + k], zkq = V[q2 * M + k];
          const double np = c * zkp - s * zkq, nq = s * zkp + c * zkq;
          changed |= bits_differ(np, zkp) || bits_differ(nq, zkq);
          V[p * M + k] = np;
          V[q2 * M + k] = nq;
        }
      }
    if (!changed) break;
  }
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < M; ++i) w[i] = A[i * M + i];
#pragma unroll
    for (int i = 0; i < M * M; ++i) Z[i] = V[i];
  }
  wave_sync();
}

__device__ __attribute__((always_inline)) void small_syev_wave(const double* Hin, int m, double* w, double* Z, double* A) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  for (int i = lane; i < m * m; i += KPC_WAVE) {
    A[i] = Hin[i];
    Z[i] = (i % m == i / m) ? 1.0 : 0.0;
  }
  wave_sync();
  for (int sweep = 0; sweep < 100; ++sweep) {
    int stop = 0;
    if (lane == 0) {
      double off = 0.0, tot = 0.0;
      for (int i = 0; i < m * m; ++i) {
        tot += A[i] * A[i];
        if (i % m != i / m) off += A[i] * A[i];
      }
      stop = off <= 1e-300 || off <= kEps * kEps * tot * 1e-4;
    }
    if (__shfl(stop, 0, KPC_WAVE)) break;
    bool changed = false;
    for (int p = 0; p < m - 1; ++p)
      for (int q2 = p + 1; q2 < m; ++q2) {
        const double apq = A[q2 * m + p];
        if (apq == 0.0) continue;
        const double app = A[p * m + p], aqq = A[q2 * m + q2];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = copysign(1.0, theta) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        wave_sync();
        for (int k = lane; k < m; k += KPC_WAVE) {
          const double akp = A[p * m + k], akq = A[q2 * m + k];
          const double np = c * akp - s * akq, nq = s * akp + c * akq;
          changed |= bits_differ(np, akp) || bits_differ(nq, akq);
          A[p * m + k] = np;
          A[q2 * m + k] = nq;
        }
        wave_sync();
        for (int k = lane; k < m; k += KPC_WAVE) {
          const double apk = A[k * m + p], aqk = A[k * m + q2];
          const double np = c * apk - s * aqk, nq = s * apk + c * aqk;
          changed |= bits_differ(np, apk) || bits_differ(nq, aqk);
          A[k * m + p] = np;
          A[k * m + q2] = nq;
        }
        for (int k = lane; k < m; k += KPC_WAVE) {
          const double zkp = Z[p * m + k], zkq = Z[q2 * m + k];
          const double np = c * zkp - s * zkq, nq = s * zkp + c * zkq;
          changed |= bits_differ(np, zkp) || bits_differ(nq, zkq);
          Z[p * m + k] = np;
          Z[q2 * m + k] = nq;
        }
        wave_sync();
      }
    if (!__any(changed)) break;
  }
  for (int i = lane; i < m; i += KPC_WAVE) w[i] = A[i * m + i];
  wave_sync();
}

// kLds: the whole working set (big arrays and derivative work) is in LDS.
// Every LDS pointer then derives from the dynamic shared array on every path,
// so the compiler addresses it as LDS (ds_read / ds_write with immediate
// offsets) instead of through generic flat pointers.
// End of an optimiser iteration (mgcv magic): the convergence test from
// iteration 4 on (score reduction and gradient norm against tol; a step
// halving that ran out counts as converged, flagged as a failed step), the
// gradient of the accepted point taken as the current one, the Hessian's
// eigen-decomposition, and the next Newton step (capped at max_newton_step;
// steepest descent when the Hessian is not positive definite) and
// steepest-descent step.  Workgroup-wide; decisions on lane 0.
__device__ __attribute__((always_inline)) void magic_conv_and_step(const MagicLds& L,
                                                                   const MagicControl& ctl, int m,
                                                                   int iter, int max_half) {
  const int tid = threadIdx.x;
  double* st = L.dctl;
  int* ic = L.ictl;
  if (tid == 0) {
    if (iter > 3) {
      double gn = 0.0;
      for (int i = 0; i < m; ++i) gn += L.grad[i] * L.grad[i];
      gn = sqrt(gn);
      const double t1 = ctl.tol * (1.0 + st[3]);
      const double t2 = pow(ctl.tol, 1.0 / 3.0) * (1.0 + fabs(st[3]));
      int conv = st[4] <= t1 && gn <= t2;
      st[5] = fmin(st[5], fmin(fabs(st[4] - t1) / t1, fabs(gn - t2) / t2));
      if (ic[6] == max_half) { conv = 1; ic[7] = 1; }
      ic[4] = conv;
    }
    for (int i = 0; i < m; ++i) L.grad[i] = L.cgrad[i];
  }
  __syncthreads();
  if (tid < KPC_WAVE) {
    switch (m) {
      case 1: small_syev_regs<1>(L.hess, L.syw, L.syZ); break;
      case 2: small_syev_regs<2>(L.hess, L.syw, L.syZ); break;
      case 3: small_syev_regs<3>(L.hess, L.syw, L.syZ); break;
      case 4: small_syev_regs<4>(L.hess, L.syw, L.syZ); break;
      default: small_syev_wave(L.hess, m, L.syw, L.syZ, L.syA); break;
    }
  }
  __syncthreads();
  if (tid == 0) {
    double* w = L.syw;
    double* Z = L.syZ;
    int sd = 0;
    for (int i = 0; i < m; ++i) {
      if (w[i] < 0.0) sd = 1;
      st[5] = fmin(st[5], fabs(w[i]) / st[3]);
    }
    ic[3] = sd;
    if (!sd) {
      for (int r = 0; r < m; ++r) {
        double v = 0.0;
        for (int c = 0; c < m; ++c) {
          double pc = 0.0;
          for (int k = 0; k < m; ++k) pc += Z[c * m + k] * L.grad[k];
          v += Z[c * m + r] * (pc / w[c]);
        }
        L.nstep[r] = -v;
      }
      double mx = fabs(L.nstep[0]);
      for (int i = 1; i < m; ++i) mx = fmax(mx, fabs(L.nstep[i]));
      if (mx > ctl.max_newton_step)
        for (int i = 0; i < m; ++i) L.nstep[i] *= ctl.max_newton_step / mx;
    }
    double mg = fabs(L.grad[0]);
    for (int i = 1; i < m; ++i) mg = fmax(mg, fabs(L.grad[i]));
    for (int i = 0; i < m; ++i) L.sstep[i] = mg == 0.0 ? 0.0 : -L.grad[i] / mg;
  }
  __syncthreads();
}

// Per-fit prologue of the GCV kernels: y into the fit's slab (and a copy),
// the timing fields cleared, y0 = Q'y, |y|^2, the starting log(sp) (the
// cached fit's when finfo is set) and the design norms.  Returns |y|^2.
__device__ __attribute__((always_inline)) double magic_prologue(const MagicLds& L, const double* D,
                                                                const DesignLayout& DL,
                                                                const double* y, int n, int q,
                                                                int m, const double* finfo) {
  const int tid = threadIdx.x;
  for (int r = tid; r < n; r += blockDim.x) { L.y[r] = y[r]; L.yorig[r] = y[r]; }
  if (tid == 0) {
    L.dctl[26] = L.dctl[27] = L.dctl[29] = 0.0;
    L.dctl[20] = L.dctl[21] = L.dctl[22] = L.dctl[23] = 0.0;
    L.dctl[8] = L.dctl[9] = L.dctl[10] = L.dctl[11] = 0.0;
    L.dctl[12] = L.dctl[13] = 0.0;
    L.ictl[28] = 0;
  }
  __syncthreads();
  // y0 = Q' y (reflectors in order)
  for (int i = 0; i < q; ++i) {
    const double* v = D + DL.QR + (size_t)i * n;
    const double tau = D[DL.tau + i];
    double part = 0.0;
    for (int r = i + tid; r < n; r += blockDim.x) part += (r == i ? 1.0 : v[r]) * L.y[r];
    const double s = block_sum(part, L.red) * tau;
    for (int r = i + tid; r < n; r += blockDim.x) L.y[r] -= s * (r == i ? 1.0 : v[r]);
    __syncthreads();
  }
  double part = 0.0;
  for (int r = tid; r < n; r += blockDim.x) part += L.yorig[r] * L.yorig[r];
  const double ynorm2 = block_sum(part, L.red);
  if (tid < m) L.lsp[tid] = finfo ? finfo[8 + tid] : D[DL.lsp0 + tid];
  for (int i = tid; i < 2 + 3 * m; i += blockDim.x) L.nrm[i] = D[DL.norms + i];
  return ynorm2;
}

// Residuals y - X beta at the point last evaluated (the selected log(sp))
// and the fit record: the cached record for a replayed fit, else status,
// score, rss, edf, iterations, evaluations, step failure, rank, log(sp),
// decision margin and the timing fields.
__device__ __attribute__((always_inline)) void magic_finish(
    const MagicLds& L, const MagicScratch& S, const double* D, const DesignLayout& DL, int n,
    int q, int m, double* ro, double* out, const double* finfo, int converged, int iter,
    long long t_start, long long t_final) {
  const int tid = threadIdx.x;
  const int* ic = L.ictl;
  const double* st = L.dctl;
  const int nr = L.ictl[0];
  const bool qr = L.ictl[13] != 0;
  if (qr && !L.ictl[15]) upper_tri_inverse_fast(L.aug, q + L.ictl[1], q, L.V, L.wk);
  double* theta = S.theta;
  for (int r = tid; r < q; r += blockDim.x) {
    double v = 0.0;
    if (qr) {
      for (int c = r; c < q; ++c) v += L.V[(size_t)c * q + r] * L.y1[c];  // R2^-1 y1
    } else {
      for (int c = 0; c < nr; ++c)
        v += L.V[(size_t)L.ord[c] * q + r] * (L.y1[c] / L.d[L.ord[c]]);
    }
    theta[r] = v;
  }
  __syncthreads();
  double* beta = S.beta;
  for (int i = tid; i < q; i += blockDim.x) beta[(int)D[DL.piv + i]] = theta[i];
  __syncthreads();
  const double* X = D + DL.X;
  for (int r = tid; r < n; r += blockDim.x) {
    double f = 0.0;
    for (int c = 0; c < q; ++c) f += X[(size_t)c * n + r] * beta[c];
    ro[r] = L.yorig[r] - f;
  }
  if (finfo) {
    for (int i = tid; i < kMagicInfo; i += blockDim.x) out[i] = finfo[i];
  } else if (tid == 0) {
    out[0] = converged ? 0.0 : -7.0;
    out[1] = L.dctl[2];
    out[2] = L.dctl[0];
    out[3] = L.dctl[1];
    out[4] = (double)iter;
    out[5] = (double)ic[9];
    out[6] = (double)ic[7];
    out[7] = (double)nr;
    for (int i = 0; i < m && i < kMagicInfo - 10; ++i) out[8 + i] = L.lsp[i];
    out[kMagicInfo - 2] = st[5];
    out[kMagicInfo - 3] = (double)(wall_clock64() - t_start);  // wall-clock ticks
    out[kMagicInfo - 4] = L.dctl[26];  // ticks in the pivoted Cholesky
    out[kMagicInfo - 5] = L.dctl[27];  // ticks in the SVD
    out[kMagicInfo - 6] = L.dctl[29];  // ticks in the derivatives
    out[kMagicInfo - 7] = (double)L.ictl[28];  // Jacobi sweeps
    if (8 + m <= 14) {  // evaluation sub-phases where the log(sp) slots leave room
      out[14] = L.dctl[8];
      out[15] = L.dctl[9];
      out[16] = L.dctl[10];
      out[17] = L.dctl[11];
      out[18] = L.dctl[12];
      out[19] = L.dctl[13];
    }
    if (8 + m <= 20) {  // phase ticks where the log(sp) slots leave room
      out[20] = L.dctl[20];  // all objective evaluations
      out[21] = L.dctl[21];  // all derivative evaluations
      out[22] = L.dctl[22];  // Q'y prologue
      out[23] = (double)(wall_clock64() - t_final);  // final fit and residuals
    }
  }
  __syncthreads();
}

template <bool kLds>
__global__ __launch_bounds__(kMagicThreads, 2) void magic_kernel(
    const MagicItem* __restrict__ items, int count, int n, int q, int m, int rk,
    const double* __restrict__ data, const double* __restrict__ design,
    double* __restrict__ resid, double* __restrict__ info, double* __restrict__ scratch,
    long long scratch_stride, int aug_in_lds, MagicControl ctl,
    const double* __restrict__ fixed_info) {
  extern __shared__ double lds_dyn[];
  const int tid = threadIdx.x;
  double* gs = scratch + (size_t)blockIdx.x * scratch_stride;
  MagicLds L = carve_magic(lds_dyn, n, q, m, kLds || aug_in_lds != 0, gs,
                           kLds || magic_deriv_in_lds(q, m));
  MagicScratch S = carve_magic_scratch(gs, q, m, rk);
  const DesignLayout DL = design_layout(n, q, m, rk);
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const MagicItem it = items[item];
    const double* D = design + (size_t)it.design_slot * DL.stride;
    const double* y = data + (size_t)it.target * n;
    double* out = info + (size_t)it.resid_slot * kMagicInfo;
    // it.pad > 0: a cached fit -- the final fit at the log(sp) of the fit
    // record fixed_info[it.pad - 1] (no optimiser), then that record restored
    const double* finfo =
        it.pad > 0 && fixed_info ? fixed_info + (size_t)(it.pad - 1) * kMagicInfo : nullptr;
    const long long t_start = wall_clock64();
    const double ynorm2 = magic_prologue(L, D, DL, y, n, q, m, finfo);
    if (tid == 0) L.dctl[22] = (double)(wall_clock64() - t_start);  // Q'y prologue
    __syncthreads();
    const int max_half = m == 1 ? ctl.max_half_single : ctl.max_half_multi;
    int iter = 0;
    if (finfo) {
      if (tid == 0) L.ictl[4] = 1;
      __syncthreads();
    } else {
    magic_eval(L, S, D, DL, n, q, m, rk, L.lsp, ynorm2, ctl.rank_tol);
    magic_derivs(L, S, D, DL, n, q, m, rk, L.lsp);
    // optimizer state (lane 0 decides, LDS broadcasts)
    double* st = L.dctl;  // st[3]=min_score st[4]=score_red
    int* ic = L.ictl;     // ic[3]=use_sd ic[4]=converged ic[5]=iter ic[6]=tries ic[7]=step_failed
    if (tid == 0) {
      st[3] = L.dctl[2];
      st[4] = 1e10;
      ic[3] = 0;
      ic[4] = 0;
      ic[7] = 0;
      ic[9] = 1;  // evaluations
      st[5] = L.dctl[6];  // smallest normalised decision margin
      for (int i = 0; i < m; ++i) { L.grad[i] = 0.0; L.nstep[i] = 0.0; L.sstep[i] = 0.0; }
    }
    __syncthreads();
    for (iter = 1; iter <= ctl.max_iterations; ++iter) {
      if (tid == 0) ic[6] = 0;
      __syncthreads();
      if (iter > 1) {
        if (tid == 0) {
          for (int i = 0; i < m; ++i) L.step[i] = ic[3] ? L.sstep[i] : L.nstep[i];
          ic[11] = 1;  // trying
        }
        __syncthreads();
        while (ic[11]) {
          if (tid == 0) {
            ic[6] += 1;
            if (ic[6] == 4 && !ic[3]) {
              ic[3] = 1;
              for (int i = 0; i < m; ++i) L.step[i] = L.sstep[i];
            }
            for (int i = 0; i < m; ++i) L.trial[i] = L.lsp[i] + L.step[i];
          }
          __syncthreads();
          magic_eval(L, S, D, DL, n, q, m, rk, L.trial, ynorm2, ctl.rank_tol);
          if (tid == 0) {
            ic[9] += 1;
            ic[12] = L.dctl[2] < st[3];
            st[5] = fmin(st[5], fmin(L.dctl[6], fabs(L.dctl[2] - st[3]) / st[3]));
          }
          __syncthreads();
          if (ic[12]) {
            if (tid == 0) {
              ic[11] = 0;
              st[4] = st[3] - L.dctl[2];
              st[3] = L.dctl[2];
              for (int i = 0; i < m; ++i) L.lsp[i] = L.trial[i];
            }
            __syncthreads();
            magic_derivs(L, S, D, DL, n, q, m, rk, L.lsp);
          } else if (tid == 0) {
            for (int i = 0; i < m; ++i) L.step[i] /= 2.0;
          }
          if (tid == 0) {
            if (ic[6] == max_half - 1 && ic[11]) for (int i = 0; i < m; ++i) L.step[i] = 0.0;
            if (ic[6] == max_half) ic[11] = 0;
          }
          __syncthreads();
        }
      }
      magic_conv_and_step(L, ctl, m, iter, max_half);
      if (ic[4]) break;
    }
    // infinity probes per coordinate
    if (ic[4]) {
      for (int c = 0; c < m; ++c) {
        const double dir = L.grad[c] < 0.0 ? 1.0 : -1.0;
        for (int pr = 1; pr <= ctl.max_probes; ++pr) {
          if (tid == 0) {
            for (int i = 0; i < m; ++i) L.trial[i] = L.lsp[i];
            L.trial[c] += dir * ctl.probe_step;
          }
          __syncthreads();
          magic_eval(L, S, D, DL, n, q, m, rk, L.trial, ynorm2, ctl.rank_tol);
          if (tid == 0) {
            ic[9] += 1;
            ic[12] = L.dctl[2] < st[3];
            st[5] = fmin(st[5], fmin(L.dctl[6], fabs(L.dctl[2] - st[3]) / st[3]));
            if (ic[12]) {
              for (int i = 0; i < m; ++i) L.lsp[i] = L.trial[i];
              st[3] = L.dctl[2];
            }
          }
          __syncthreads();
          if (!ic[12]) break;
        }
      }
    }
    }  // optimiser (skipped for a cached fit)
    const int converged = L.ictl[4];
    // final fit at the selected log(sp)
    const long long t_final = wall_clock64();
    magic_eval(L, S, D, DL, n, q, m, rk, L.lsp, ynorm2, ctl.rank_tol);
    magic_finish(L, S, D, DL, n, q, m, resid + (size_t)it.resid_slot * n, out, finfo, converged,
                 iter, t_start, t_final);
  }
}

// ---------------------------------------------------------------- slots
// magic for additive designs with m = K rank-8 penalty blocks (q = 1 + 9 K,
// K = 3 or 4):
// the optimiser of magic_kernel with every score routed through one site.
// The waves of the workgroup score up to nslot points of the current
// sequence at once -- one iteration's step-halving trials, one coordinate's
// infinity probes, or a single point -- on the one-wave certified-QR path
// (wave_eval_additive, the arithmetic of magic_eval's QR path); a point not
// certified for that path is scored by magic_eval on the whole workgroup.
// The points of a sequence are fixed before any is scored (try t uses the
// step the serial loop reaches after t - 1 failures; probe p the point after
// p - 1 accepted probes) and are consumed strictly in order by the serial
// accept / count / margin rules, so the trajectory, the evaluation count and
// every output equal the one-at-a-time optimiser's.
enum : int { kPhFirst = 0, kPhHalve = 1, kPhProbe = 2, kPhRecomp = 3, kPhFinal = 4, kPhDone = 5 };
// Control words in ictl beyond magic_kernel's: [16] phase, [17] pending work
// (1: derivatives at lsp, 2: convergence test and next step), [18] points in
// the sequence, [19] next point, [20] probe coordinate, [22] point for the
// workgroup path (-1: none), [23] the try accepted at the last iteration.

// Probe sequence of coordinate c (the final fit once every coordinate is done).
__device__ __forceinline__ void slot_start_probe(const MagicLds& L, const MagicControl& ctl, int m,
                                                 int c) {
  int* ic = L.ictl;
  if (c >= m || ctl.max_probes <= 0) {
    for (int i = 0; i < m; ++i) L.trials[i] = L.lsp[i];
    ic[16] = kPhFinal;
    ic[18] = 1;
    ic[19] = 0;
    return;
  }
  const double dir = L.grad[c] < 0.0 ? 1.0 : -1.0;
  for (int pr = 0; pr < ctl.max_probes; ++pr) {
    double* t = L.trials + (size_t)pr * m;
    for (int i = 0; i < m; ++i) t[i] = L.lsp[i];
    t[c] = (pr == 0 ? L.lsp[c] : L.trials[(size_t)(pr - 1) * m + c]) + dir * ctl.probe_step;
  }
  ic[16] = kPhProbe;
  ic[18] = ctl.max_probes;
  ic[19] = 0;
  ic[20] = c;
}

// Step-halving trials of the next iteration in try order.
__device__ __forceinline__ void slot_start_halving(const MagicLds& L, int m, int max_half) {
  int* ic = L.ictl;
  double* sv = L.step;
  int sd = ic[3];
  for (int i = 0; i < m; ++i) sv[i] = sd ? L.sstep[i] : L.nstep[i];
  for (int t = 1; t <= max_half; ++t) {
    if (t == 4 && !sd) {
      sd = 1;
      for (int i = 0; i < m; ++i) sv[i] = L.sstep[i];
    }
    for (int i = 0; i < m; ++i) L.trials[(size_t)(t - 1) * m + i] = L.lsp[i] + sv[i];
    for (int i = 0; i < m; ++i) sv[i] /= 2.0;
    if (t == max_half - 1)
      for (int i = 0; i < m; ++i) sv[i] = 0.0;
  }
  ic[16] = kPhHalve;
  ic[18] = max_half;
  ic[19] = 0;
}

// Slot 0 scored its point into the workgroup's buffers: its results become
// the workgroup's last evaluation (as magic_eval would leave them).
__device__ __forceinline__ void slot_load0(const MagicLds& L, int q, int mr, const double* r) {
  L.ictl[0] = q;
  L.ictl[1] = mr;
  L.ictl[13] = 1;
  L.ictl[15] = 0;
  L.ictl[2] = 0;
  L.dctl[0] = r[1];
  L.dctl[1] = r[2];
  L.dctl[2] = r[0];
  L.dctl[6] = r[3];
}

// Lane 0: the score of point k of the current sequence; in_l when the point's
// factors are in the workgroup's buffers.  False when the sequence ends here.
__device__ __forceinline__ bool slot_consume(const MagicLds& L, const MagicControl& ctl, int m,
                                             int max_half, int k, double score, double margin,
                                             bool in_l) {
  int* ic = L.ictl;
  double* st = L.dctl;  // st[3] min score, st[4] score reduction, st[5] margin
  switch (ic[16]) {
    case kPhFirst:
      st[3] = score;
      st[4] = 1e10;
      ic[3] = ic[4] = ic[7] = 0;
      ic[9] = 1;
      st[5] = margin;
      for (int i = 0; i < m; ++i) L.grad[i] = L.nstep[i] = L.sstep[i] = 0.0;
      ic[5] = 1;
      ic[6] = 0;
      ic[17] = 3;
      return false;
    case kPhHalve: {
      const int t = k + 1;
      ic[6] = t;
      ic[9] += 1;
      ic[19] = t;
      st[5] = fmin(st[5], fmin(margin, fabs(score - st[3]) / st[3]));
      if (score < st[3]) {
        st[4] = st[3] - score;
        st[3] = score;
        ic[23] = k;  // the next iteration's predicted accepted try
        for (int i = 0; i < m; ++i) L.lsp[i] = L.trials[(size_t)k * m + i];
        if (in_l) {
          ic[17] = 3;
        } else {
          // re-score the accepted point into the workgroup's buffers first
          for (int i = 0; i < m; ++i) L.trials[i] = L.lsp[i];
          ic[16] = kPhRecomp;
          ic[18] = 1;
          ic[19] = 0;
        }
        return false;
      }
      if (t == max_half) {
        ic[17] = 2;
        return false;
      }
      return true;
    }
    case kPhRecomp:
      ic[17] = 3;
      return false;
    case kPhProbe: {
      ic[9] += 1;
      ic[19] = k + 1;
      const bool acc = score < st[3];
      st[5] = fmin(st[5], fmin(margin, fabs(score - st[3]) / st[3]));
      if (acc) {
        for (int i = 0; i < m; ++i) L.lsp[i] = L.trials[(size_t)k * m + i];
        st[3] = score;
        if (k + 1 < ic[18]) return true;
      }
      slot_start_probe(L, ctl, m, ic[20] + 1);
      return false;
    }
    default:  // kPhFinal
      ic[16] = kPhDone;
      return false;
  }
}

template <int K>
__global__ __launch_bounds__(kMagicThreads, K == 3 ? 1 : 2) void magic_slots_kernel(
    const MagicItem* __restrict__ items, int count, int n, const double* __restrict__ data,
    const double* __restrict__ design, double* __restrict__ resid, double* __restrict__ info,
    double* __restrict__ scratch, long long scratch_stride, MagicControl ctl,
    const double* __restrict__ fixed_info, int nslot) {
  constexpr int m = K, rk = 8, q = 1 + 9 * K, mr = 8 * K;
  extern __shared__ double lds_dyn[];
  const int tid = threadIdx.x;
  const int wid = tid / KPC_WAVE;
  double* gs = scratch + (size_t)blockIdx.x * scratch_stride;
  const MagicLds L = carve_magic(lds_dyn, n, q, m, true, gs, true, nslot);
  const MagicScratch S = carve_magic_scratch(gs, q, m, rk);
  const DesignLayout DL = design_layout(n, q, m, rk);
  const int max_half = ctl.max_half_multi;
  int* ic = L.ictl;
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const MagicItem it = items[item];
    const double* D = design + (size_t)it.design_slot * DL.stride;
    const double* y = data + (size_t)it.target * n;
    double* out = info + (size_t)it.resid_slot * kMagicInfo;
    const double* finfo =
        it.pad > 0 && fixed_info ? fixed_info + (size_t)(it.pad - 1) * kMagicInfo : nullptr;
    const long long t_start = wall_clock64();
    const double ynorm2 = magic_prologue(L, D, DL, y, n, q, m, finfo);
    // the fit's constant inputs of the one-wave path, staged once
    for (int e = tid; e < q * q; e += blockDim.x) L.Rst[e] = D[DL.R + e];
    for (int e = tid; e < mr * q; e += blockDim.x) L.rSst[e] = D[DL.rS + e];
    for (int r = tid; r < q; r += blockDim.x) L.yst[r] = L.y[r];
    if (tid == 0) {
      L.dctl[22] = (double)(wall_clock64() - t_start);  // Q'y prologue
      for (int i = 0; i < m; ++i) L.trials[i] = L.lsp[i];
      ic[16] = finfo ? kPhFinal : kPhFirst;
      ic[18] = 1;
      ic[19] = 0;
      ic[17] = 0;
      ic[4] = finfo ? 1 : 0;
      ic[5] = 0;
      ic[23] = 0;
    }
    __syncthreads();
    // every round consumes a point or ends a sequence; the bound is a guard
    const int max_rounds = 4 * (ctl.max_iterations * (max_half + 2) + m * ctl.max_probes + 8);
    for (int round = 0; ic[16] != kPhDone && round < max_rounds; ++round) {
      const int phase = ic[16];
      const int base = ic[19];
      const int cnt =
          min(phase == kPhHalve || phase == kPhProbe ? nslot : 1, ic[18] - base);
      // Point pi of the round is scored into the workgroup's buffers by wave 0:
      // in a halving sequence the try accepted at the last iteration (the
      // likeliest to be accepted again, so its derivatives need no re-score);
      // the other points, in order, go to waves 1-3 and a fifth to wave 0
      // beside pi.  Point s's results are at sres[8 s] whichever wave scored it.
      const int pred = phase == kPhHalve ? ic[23] - base : 0;
      const int pi = pred < 0 ? 0 : (pred >= cnt ? cnt - 1 : pred);
      const long long tr0 = wall_clock64();
      if (wid == 0) {
        const double* ta = L.trials + (size_t)(base + pi) * m;
        bool two = false;
        if constexpr (K == 3) {
          two = cnt >= 5;
          if (two) {
            const int sb = 3 < pi ? 3 : 4;  // the fourth point other than pi
            wave_eval_additive2<K>(L.Rst, L.rSst, L.yst, L.nrm, ta,
                                   L.trials + (size_t)(base + sb) * m, n, ynorm2, ctl.rank_tol,
                                   L.aug, q + mr, L.Qm, q + mr, L.y1, L.sres + 8 * pi,
                                   L.dctl + 8, L.sQ4, q, L.sy1, L.sres + 8 * sb);
          }
        }
        if (!two)
          wave_eval_additive<K>(L.Rst, L.rSst, L.yst, L.nrm, ta, n, ynorm2, ctl.rank_tol, L.aug,
                                q + mr, L.Qm, q + mr, L.y1, L.sres + 8 * pi, L.dctl + 8);
      } else if (wid < cnt) {
        const int o = wid - 1;               // this wave's rank among the other points
        const int sp = o < pi ? o : o + 1;   // its point
        double* Q1 = wid == 1 ? L.X4 : wid == 2 ? L.V : L.sQ3;
        wave_eval_additive<K>(L.Rst, L.rSst, L.yst, L.nrm, L.trials + (size_t)(base + sp) * m,
                              n, ynorm2, ctl.rank_tol, nullptr, q + mr, Q1, q,
                              L.sy1 + (size_t)wid * q, L.sres + 8 * sp, nullptr);
      }
      __syncthreads();
      if (tid == 0) {
        L.dctl[20] += (double)(wall_clock64() - tr0);
        ic[17] = 0;
        ic[22] = -1;
        for (int s = 0; s < cnt; ++s) {
          const double* r = L.sres + 8 * s;
          if (r[4] == 0.0) {
            ic[22] = base + s;
            break;
          }
          if (s == pi) slot_load0(L, q, mr, r);
          if (!slot_consume(L, ctl, m, max_half, base + s, r[0], r[3], s == pi)) break;
        }
      }
      __syncthreads();
      if (ic[22] >= 0) {
        const int k = ic[22];
        magic_eval(L, S, D, DL, n, q, m, rk, L.trials + (size_t)k * m, ynorm2, ctl.rank_tol);
        if (tid == 0) slot_consume(L, ctl, m, max_half, k, L.dctl[2], L.dctl[6], true);
        __syncthreads();
      }
      const int work = ic[17];
      if (work & 1) magic_derivs(L, S, D, DL, n, q, m, rk, L.lsp);
      if (work & 2) {
        magic_conv_and_step(L, ctl, m, ic[5], max_half);
        if (tid == 0) {
          if (ic[4]) {
            slot_start_probe(L, ctl, m, 0);
          } else if (ic[5] + 1 > ctl.max_iterations) {
            ic[5] += 1;
            slot_start_probe(L, ctl, m, m);  // out of iterations: the final fit
          } else {
            ic[5] += 1;
            ic[6] = 0;
            slot_start_halving(L, m, max_half);
          }
        }
        __syncthreads();
      }
    }
    const long long t_final = wall_clock64();
    magic_finish(L, S, D, DL, n, q, m, resid + (size_t)it.resid_slot * n, out, finfo, ic[4], ic[5],
                 t_start, t_final);
  }
}

size_t magic_lds_bytes(int n, int q, int m, bool aug_in_lds) {
  return magic_lds_doubles(n, q, m, aug_in_lds) * sizeof(double);
}

bool magic_aug_fits_lds(int n, int q, int m) {
  return magic_lds_bytes(n, q, m, true) <= 150 * 1024;
}

long long geometry_scratch_doubles(int q) {
  long long s = 29LL * 29 + 4 * 29 + 64 + 2LL * q + 64 + (long long)q * q + 2LL * q;
  return (s + 31) / 32 * 32;
}

hipError_t launch_geometry(const GeometryItem* items, int count, int n, int q, int m, int dim,
                           int rk, const double* smooth_X, const double* smooth_S, double* design,
                           double* status_out, long long smooth_x_stride, double* scratch,
                           int grid, hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  // the QR in LDS when X (n x q) fits beside the kernel's static LDS in a CU
  hipFuncAttributes fa;
  hipError_t err = hipFuncGetAttributes(&fa, (const void*)geometry_kernel<true>);
  if (err != hipSuccess) return err;
  const size_t dyn = (size_t)n * q * sizeof(double);
  if (dyn + fa.sharedSizeBytes <= 163840) {
    err = hipFuncSetAttribute((const void*)geometry_kernel<true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(geometry_kernel<true>, dim3(grid), dim3(kGeomThreads), dyn, stream, items,
                       count, n, q, m, dim, rk, smooth_X, smooth_S, design, status_out,
                       smooth_x_stride, scratch, geometry_scratch_doubles(q));
  } else {
    hipLaunchKernelGGL(geometry_kernel<false>, dim3(grid), dim3(kGeomThreads), 0, stream, items,
                       count, n, q, m, dim, rk, smooth_X, smooth_S, design, status_out,
                       smooth_x_stride, scratch, geometry_scratch_doubles(q));
  }
  return hipGetLastError();
}

// Trial slots per launch: KPC_MAGIC_SLOTS=1 scores step-halving trials one
// at a time on the workgroup path (outputs are the same either way); read at
// every launch.
static int magic_slot_limit() {
  const char* e = getenv("KPC_MAGIC_SLOTS");
  const int v = e ? atoi(e) : 5;
  return v < 1 ? 1 : v > 5 ? 5 : v;
}

hipError_t launch_magic(const MagicItem* items, int count, int n, int q, int m, int rk,
                        const double* data, const double* design, double* resid, double* info,
                        double* scratch, int grid, const MagicControl& ctl, hipStream_t stream,
                        const double* fixed_info) {
  if (count <= 0) return hipSuccess;
  const bool in_lds = magic_aug_fits_lds(n, q, m);
  const bool all_lds = in_lds && magic_deriv_in_lds(q, m);
  const int nslot = all_lds && ctl.max_half_multi >= 1 && ctl.max_half_multi <= kSpecTries &&
                            ctl.max_probes <= kSpecTries
                        ? std::min(magic_spec_slots(n, q, m, rk), magic_slot_limit())
                        : 1;
  const long long stride = magic_scratch_doubles(q, m, rk, n);
  if (nslot > 1) {
    const size_t lds = magic_lds_doubles(n, q, m, true, nslot) * sizeof(double);
    const void* fn =
        m == 3 ? (const void*)magic_slots_kernel<3> : (const void*)magic_slots_kernel<4>;
    hipError_t err = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (err != hipSuccess) return err;
    if (m == 3)
      hipLaunchKernelGGL(magic_slots_kernel<3>, dim3(grid), dim3(256), lds, stream, items, count, n,
                         data, design, resid, info, scratch, stride, ctl, fixed_info, nslot);
    else
      hipLaunchKernelGGL(magic_slots_kernel<4>, dim3(grid), dim3(256), lds, stream, items, count, n,
                         data, design, resid, info, scratch, stride, ctl, fixed_info, nslot);
    return hipGetLastError();
  }
  const size_t lds = magic_lds_bytes(n, q, m, in_lds);
  const void* fn = all_lds ? (const void*)magic_kernel<true> : (const void*)magic_kernel<false>;
  hipError_t err = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (err != hipSuccess) return err;
  if (all_lds)
    hipLaunchKernelGGL(magic_kernel<true>, dim3(grid), dim3(magic_threads(q)), lds, stream, items,
                       count, n, q, m, rk, data, design, resid, info, scratch, stride, 1, ctl,
                       fixed_info);
  else
    hipLaunchKernelGGL(magic_kernel<false>, dim3(grid), dim3(magic_threads(q)), lds, stream, items,
                       count, n, q, m, rk, data, design, resid, info, scratch, stride,
                       in_lds ? 1 : 0, ctl, fixed_info);
  return hipGetLastError();
}

}  // namespace kpc

namespace kpc {

// GCV score, gradient and Hessian at caller-supplied log(sp): the hand-written
// derivative of the magic objective, exposed on its own for testing against
// central differences and for external Newton drivers.
template <bool kLds>
__global__ __launch_bounds__(kMagicThreads, 2) void gcv_eval_kernel(
    const MagicItem* __restrict__ items, int count, int n, int q, int m, int rk,
    const double* __restrict__ data, const double* __restrict__ design,
    const double* __restrict__ lsp_in, double* __restrict__ out, double* __restrict__ scratch,
    long long scratch_stride, int aug_in_lds, double rank_tol) {
  extern __shared__ double lds_dyn[];
  const int tid = threadIdx.x;
  double* gs = scratch + (size_t)blockIdx.x * scratch_stride;
  MagicLds L = carve_magic(lds_dyn, n, q, m, kLds || aug_in_lds != 0, gs,
                           kLds || magic_deriv_in_lds(q, m));
  MagicScratch S = carve_magic_scratch(gs, q, m, rk);
  const DesignLayout DL = design_layout(n, q, m, rk);
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const MagicItem it = items[item];
    const double* D = design + (size_t)it.design_slot * DL.stride;
    const double* y = data + (size_t)it.target * n;
    for (int r = tid; r < n; r += blockDim.x) { L.y[r] = y[r]; L.yorig[r] = y[r]; }
    __syncthreads();
    for (int i = 0; i < q; ++i) {
      const double* v = D + DL.QR + (size_t)i * n;
      const double tau = D[DL.tau + i];
      double part = 0.0;
      for (int r = i + tid; r < n; r += blockDim.x) part += (r == i ? 1.0 : v[r]) * L.y[r];
      const double s = block_sum(part, L.red) * tau;
      for (int r = i + tid; r < n; r += blockDim.x) L.y[r] -= s * (r == i ? 1.0 : v[r]);
      __syncthreads();
    }
    double part = 0.0;
    for (int r = tid; r < n; r += blockDim.x) part += L.yorig[r] * L.yorig[r];
    const double ynorm2 = block_sum(part, L.red);
    if (tid < m) L.lsp[tid] = lsp_in[(size_t)item * m + tid];
    for (int i = tid; i < 2 + 3 * m; i += blockDim.x) L.nrm[i] = D[DL.norms + i];
    __syncthreads();
    magic_eval(L, S, D, DL, n, q, m, rk, L.lsp, ynorm2, rank_tol);
    magic_derivs(L, S, D, DL, n, q, m, rk, L.lsp);
    double* o = out + (size_t)item * (1 + m + m * m);
    if (tid == 0) {
      o[0] = L.dctl[2];
      for (int i = 0; i < m; ++i) o[1 + i] = L.cgrad[i];
      for (int i = 0; i < m * m; ++i) o[1 + m + i] = L.hess[i];
    }
    __syncthreads();
  }
}

hipError_t launch_gcv_eval(const MagicItem* items, int count, int n, int q, int m, int rk,
                           const double* data, const double* design, const double* lsp,
                           double* out, double* scratch, int grid, double rank_tol,
                           hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  const bool in_lds = magic_aug_fits_lds(n, q, m);
  const size_t lds = magic_lds_bytes(n, q, m, in_lds);
  const bool all_lds = in_lds && magic_deriv_in_lds(q, m);
  const void* fn = all_lds ? (const void*)gcv_eval_kernel<true> : (const void*)gcv_eval_kernel<false>;
  hipError_t err = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (err != hipSuccess) return err;
  if (all_lds)
    hipLaunchKernelGGL(gcv_eval_kernel<true>, dim3(grid), dim3(magic_threads(q)), lds, stream, items,
                       count, n, q, m, rk, data, design, lsp, out, scratch,
                       magic_scratch_doubles(q, m, rk, n), 1, rank_tol);
  else
    hipLaunchKernelGGL(gcv_eval_kernel<false>, dim3(grid), dim3(magic_threads(q)), lds, stream, items,
                       count, n, q, m, rk, data, design, lsp, out, scratch,
                       magic_scratch_doubles(q, m, rk, n), in_lds ? 1 : 0, rank_tol);
  return hipGetLastError();
}

}  // namespace kpc
