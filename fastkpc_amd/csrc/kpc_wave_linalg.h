// Wave-synchronous dense kernels for the GCV evaluation (one wave64 does the
// whole factorisation; lanes communicate through LDS and cross-lane moves,
// with no workgroup barrier on the critical path).
//
//   pstf2_upper_wave  LAPACK dpstf2 (upper, default tolerance) semantics:
//                     first-maximum pivoting, the three disjoint swaps, stop
//                     when the pivot is <= q eps max(diag) or NaN.
//   jacobi_svd_wave   one-sided Jacobi SVD with a round-robin parallel
//                     ordering: column pairs are spread over lane groups, the
//                     rotation of a pair is computed redundantly by its lanes,
//                     V accumulates the rotations; sweeps stop when no pair
//                     needed a rotation (|a'b| <= eps sqrt(a'a b'b)).
#pragma once
#include "kpc_common.h"

namespace kpc {

// Orders LDS traffic between the lanes of one wave (no instruction is
// emitted: LDS operations of a wave execute in order; the fences stop the
// compiler from moving memory operations across this point).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// 1/x and 1/sqrt(x) from the hardware estimates plus two Newton steps
// (quadratic convergence from the estimate's ~2^-22 to full precision).
__device__ __forceinline__ double nrcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

__device__ __forceinline__ double nrsq(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  y = y * fma(-h * y, y, 1.5);
  return y;
}

// Sum over aligned groups of `width` lanes (width a power of two <= 64).
__device__ __forceinline__ double group_sum(double v, int width) {
  for (int off = width >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, KPC_WAVE);
  return v;
}

// First maximum (largest value, smallest index on ties) over the wave.
__device__ __forceinline__ void wave_argmax_first(double& v, int& idx) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off, KPC_WAVE);
    const int oi = __shfl_xor(idx, off, KPC_WAVE);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
}

// Pivoted Cholesky (dpstf2, upper) of the q x q symmetric A (column-major,
// ld q) by one wave; work: 2q doubles.  Returns the rank; piv receives the
// pivot order.  Must be called by all lanes of exactly one wave.
__device__ int pstf2_upper_wave(double* A, int q, int* piv, double* work) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  for (int i = lane; i < q; i += KPC_WAVE) {
    piv[i] = i;
    work[i] = 0.0;
  }
  double bv = -__builtin_inf();
  int bi = 0x7fffffff;
  for (int i = lane; i < q; i += KPC_WAVE) {
    const double v = A[(size_t)i * q + i];
    if (v > bv || (v == bv && i < bi)) {
      bv = v;
      bi = i;
    }
  }
  wave_argmax_first(bv, bi);
  wave_sync();
  if (bv <= 0.0 || isnan(bv)) return 0;
  const double dstop = (double)q * 2.220446049250313e-16 * bv;
  int rank = q;
  for (int j = 0; j < q; ++j) {
    double ajj = -__builtin_inf();
    int pvt = 0x7fffffff;
    for (int i = j + lane; i < q; i += KPC_WAVE) {
      if (j > 0) work[i] += A[(size_t)i * q + j - 1] * A[(size_t)i * q + j - 1];
      const double w2 = A[(size_t)i * q + i] - work[i];
      work[q + i] = w2;
      if (w2 > ajj || (w2 == ajj && i < pvt)) {
        ajj = w2;
        pvt = i;
      }
    }
    wave_argmax_first(ajj, pvt);
    if (j == 0) {
      ajj = bv;
      pvt = bi;
    }
    wave_sync();
    if (j > 0 && (ajj <= dstop || isnan(ajj))) {
      if (lane == 0) A[(size_t)j * q + j] = ajj;
      rank = j;
      break;
    }
    if (j != pvt) {
      for (int k = lane; k < j; k += KPC_WAVE) {
        const double t = A[(size_t)j * q + k];
        A[(size_t)j * q + k] = A[(size_t)pvt * q + k];
        A[(size_t)pvt * q + k] = t;
      }
      for (int k = pvt + 1 + lane; k < q; k += KPC_WAVE) {
        const double t = A[(size_t)k * q + j];
        A[(size_t)k * q + j] = A[(size_t)k * q + pvt];
        A[(size_t)k * q + pvt] = t;
      }
      for (int k = j + 1 + lane; k < pvt; k += KPC_WAVE) {
        const double t = A[(size_t)k * q + j];
        A[(size_t)k * q + j] = A[(size_t)pvt * q + k];
        A[(size_t)pvt * q + k] = t;
      }
      if (lane == 0) {
        A[(size_t)pvt * q + pvt] = A[(size_t)j * q + j];
        const double t = work[j];
        work[j] = work[pvt];
        work[pvt] = t;
        const int it = piv[pvt];
        piv[pvt] = piv[j];
        piv[j] = it;
      }
      wave_sync();
    }
    const double rjj = sqrt(ajj);
    const double rinv = 1.0 / rjj;
    for (int c = j + 1 + lane; c < q; c += KPC_WAVE) {
      double s = 0.0;
      for (int k = 0; k < j; ++k) s += A[(size_t)c * q + k] * A[(size_t)j * q + k];
      A[(size_t)c * q + j] = (A[(size_t)c * q + j] - s) * rinv;
    }
    wave_sync();
    if (lane == 0) A[(size_t)j * q + j] = rjj;
    wave_sync();
  }
  return rank;
}

// One-sided Jacobi SVD of the rows x q matrix A (column-major, ld rows) by one
// wave.  On return the columns of A are U * diag(d) normalised to U, d holds
// the singular values by column, ord the descending order (ties keep column
// order) and V (q x q, column-major) the accumulated rotations.  Returns the
// number of sweeps.
__device__ int jacobi_svd_wave(double* A, int rows, int q, double* V, double* d, int* ord) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  constexpr double kEps = 2.220446049250313e-16;
  for (int e = lane; e < q * q; e += KPC_WAVE) V[e] = (e % q == e / q) ? 1.0 : 0.0;
  const int qe = q + (q & 1);
  const int npairs = qe / 2;
  const int qm = qe - 1;
  int lpp = 64;
  while (lpp > 1 && npairs * lpp > KPC_WAVE) lpp >>= 1;
  const int groups = KPC_WAVE / lpp;
  const int sub = lane % lpp;
  const int g0 = lane / lpp;
  wave_sync();
  int sweep = 0;
  for (; sweep < 60; ++sweep) {
    int rotated = 0;
    for (int rnd = 0; rnd < qm; ++rnd) {
      for (int pair = g0; pair < npairs + (groups - npairs % groups) % groups; pair += groups) {
        // round-robin: pair 0 meets the fixed column qe-1, pair p > 0 pairs
        // (rnd + p) and (rnd - p) mod (qe - 1)
        int a = -1, b = -1;
        if (pair < npairs) {
          if (pair == 0) {
            a = rnd;
            b = qe - 1;
          } else {
            a = (rnd + pair) % qm;
            b = (rnd - pair + qm) % qm;
          }
          if (a > b) {
            const int t = a;
            a = b;
            b = t;
          }
        }
        const bool active = pair < npairs && b < q;
        double al = 0.0, be = 0.0, ga = 0.0;
        const double* ca = A + (size_t)(active ? a : 0) * rows;
        const double* cb = A + (size_t)(active ? b : 0) * rows;
        if (active)
          for (int r = sub; r < rows; r += lpp) {
            const double x = ca[r], y = cb[r];
            al += x * x;
            be += y * y;
            ga += x * y;
          }
        al = group_sum(al, lpp);
        be = group_sum(be, lpp);
        ga = group_sum(ga, lpp);
        if (active && ga != 0.0 && ga * ga > kEps * kEps * (al * be)) {
          const double zeta = (be - al) * (0.5 * nrcp(ga));
          double t;
          if (fabs(zeta) > 1e100) {
            t = 0.5 / zeta;
          } else {
            const double z1 = fma(zeta, zeta, 1.0);
            const double root = z1 * nrsq(z1);
            t = copysign(nrcp(fabs(zeta) + root), zeta);
          }
          const double c = nrsq(fma(t, t, 1.0));
          const double sn = c * t;
          double* wa = A + (size_t)a * rows;
          double* wb = A + (size_t)b * rows;
          for (int r = sub; r < rows; r += lpp) {
            const double x = wa[r], y = wb[r];
            wa[r] = c * x - sn * y;
            wb[r] = sn * x + c * y;
          }
          double* va = V + (size_t)a * q;
          double* vb = V + (size_t)b * q;
          for (int r = sub; r < q; r += lpp) {
            const double x = va[r], y = vb[r];
            va[r] = c * x - sn * y;
            vb[r] = sn * x + c * y;
          }
          rotated = 1;
        }
      }
      wave_sync();
    }
    if (!__any(rotated)) {
      ++sweep;
      break;
    }
  }
  // singular values, normalised columns, descending order (stable)
  for (int c = 0; c < q; ++c) {
    double* col = A + (size_t)c * rows;
    double s = 0.0;
    for (int r = lane; r < rows; r += KPC_WAVE) s += col[r] * col[r];
    s = sqrt(group_sum(s, KPC_WAVE));
    if (lane == 0) d[c] = s;
    const double inv = s > 0.0 ? 1.0 / s : 0.0;
    for (int r = lane; r < rows; r += KPC_WAVE) col[r] *= inv;
  }
  wave_sync();
  for (int i = lane; i < q; i += KPC_WAVE) {
    const double di = d[i];
    int r = 0;
    for (int j = 0; j < q; ++j) r += (d[j] > di) || (d[j] == di && j < i);
    ord[r] = i;
  }
  wave_sync();
  return sweep;
}

}  // namespace kpc
