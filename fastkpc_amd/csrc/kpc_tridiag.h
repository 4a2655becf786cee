// Symmetric tridiagonal eigen-solvers shared by the Lanczos kernels (dcov
// components, thin-plate smooth setup): Sturm counts, multisection for
// selected eigenvalues, twisted factorisations for their eigenvectors, and the
// Ritz vectors Q S on FP64 matrix cores.
// Dependent FP64 operations cost tens of cycles each on CDNA4, so the serial
// recurrences use hardware reciprocal estimates refined by Newton steps.
#pragma once
#include "kpc_common.h"

namespace kpc {

// 1/x from the hardware estimate and one (Sturm counts) or two Newton steps.
__device__ __forceinline__ double lz_rcp1(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
}
__device__ __forceinline__ double lz_rcp(double x) {
  const double r = lz_rcp1(x);
  return fma(r, fma(-x, r, 1.0), r);
}

// Number of eigenvalues of the tridiagonal (d, e2 = e^2) strictly below x.
// Four rows per trip: their loads and shifts d_i - x are issued ahead of the
// pivot chain, which then waits on no memory (same operations per row).
__device__ inline int lz_sturm(const double* d, const double* e2, int n, double x, double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  cnt += q < 0.0;
  int i = 1;
  for (; i + 4 <= n; i += 4) {
    double dx[4], ev[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      dx[u] = d[i + u] - x;
      ev[u] = e2[i - 1 + u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      q = dx[u] - ev[u] * lz_rcp1(q);
      if (fabs(q) < pivmin) q = -pivmin;
      cnt += q < 0.0;
    }
  }
  for (; i < n; ++i) {
    q = (d[i] - x) - e2[i - 1] * lz_rcp1(q);
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0.0;
  }
  return cnt;
}

// Eigenvector of the tridiagonal (al, be) for the eigenvalue th by a twisted
// factorization (as in LAPACK dlar1v / MRRR): forward pivots d_i of
// T - th I = L D L', backward pivots r_i of U R U', twist index
// k = argmin |gamma_k| with gamma_k = d_k + r_k - (al_k - th); then x_k = 1,
// x_i = -be_i x_{i+1} / d_i upward and x_{i+1} = -be_i x_i / r_{i+1} downward,
// which satisfies every row except k (residual gamma_k).  rd, rr: m doubles
// of workspace (reciprocal pivots).  Writes the unit vector to z when z is
// non-null; returns its last component.
__device__ inline double lz_twisted(const double* al, const double* be, int m, double th, double pivmin,
                             double* rd, double* rr, double* z) {
  // forward and backward pivot sweeps, interleaved for ILP; four steps per
  // trip with their diagonal shifts and squared off-diagonals loaded first
  // (the pivot stores to LDS would otherwise keep later loads behind them)
  double d = al[0] - th;
  double r = al[m - 1] - th;
  int i = 0;
  for (; i + 4 < m; i += 4) {
    double fa[4], fb[4], ba[4], bb[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      fa[u] = al[i + u + 1] - th;
      fb[u] = be[i + u] * be[i + u];
      ba[u] = al[m - 2 - i - u] - th;
      bb[u] = be[m - 2 - i - u] * be[m - 2 - i - u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (fabs(d) < pivmin) d = -pivmin;
      if (fabs(r) < pivmin) r = -pivmin;
      const double id = lz_rcp(d), ir = lz_rcp(r);
      rd[i + u] = id;
      rr[m - 1 - i - u] = ir;
      d = fa[u] - fb[u] * id;
      r = ba[u] - bb[u] * ir;
    }
  }
  for (; i < m; ++i) {
    if (fabs(d) < pivmin) d = -pivmin;
    if (fabs(r) < pivmin) r = -pivmin;
    const double id = lz_rcp(d), ir = lz_rcp(r);
    rd[i] = id;
    rr[m - 1 - i] = ir;
    if (i + 1 < m) {
      d = (al[i + 1] - th) - be[i] * be[i] * id;
      r = (al[m - 2 - i] - th) - be[m - 2 - i] * be[m - 2 - i] * ir;
    }
  }
  // twist: |gamma_k| = |1/rd_k + 1/rr_k - (al_k - th)|
  int kt = 0;
  double gmin = 1e308;
  for (int i = 0; i < m; ++i) {
    const double g = fabs(lz_rcp(rd[i]) + lz_rcp(rr[i]) - (al[i] - th));
    if (g < gmin) {
      gmin = g;
      kt = i;
    }
  }
  double n0 = 1.0, n1 = 0.0, x = 1.0, last = kt == m - 1 ? 1.0 : 0.0;
  for (int i = kt - 1; i >= 0; --i) {
    x = -be[i] * x * rd[i];
    n1 += x * x;
  }
  x = 1.0;
  for (int i = kt; i < m - 1; ++i) {
    x = -be[i] * x * rr[i + 1];
    n0 += x * x;
  }
  if (kt < m - 1) last = x;
  const double inv = 1.0 / sqrt(n0 + n1);
  if (z) {
    // the two recurrences again, scaled on the way out (same products, so the
    // same values as storing them first and scaling in place, without reading
    // z back)
    z[kt] = inv;
    x = 1.0;
    for (int i = kt - 1; i >= 0; --i) {
      x = -be[i] * x * rd[i];
      z[i] = x * inv;
    }
    x = 1.0;
    for (int i = kt; i < m - 1; ++i) {
      x = -be[i] * x * rr[i + 1];
      z[i + 1] = x * inv;
    }
  }
  return last * inv;
}

// Sturm multisection for the candidate eigenvalues of the tridiagonal
// (al, e2): candidate c has ascending index idx(c) = c < nlow ? c : hs + c - nlow.
// G lanes per candidate probe G interior points of its bracket per round (the
// bracket shrinks by G + 1); rounds stop when every bracket is below
// rtol max(|lo|, |hi|) + atol + pivmin (Sturm counts are exact only for a
// perturbation of T of order eps |T|, so atol ~ eps |T| costs no accuracy and
// bounds the rounds for eigenvalues near zero).  Brackets persist in clo/chi
// (init: start from [lo0, hi0]); cval receives the midpoints.  All threads
// must call it.
__device__ inline void lz_multisect(const double* al, const double* e2, int m, double pivmin, int ncand,
                             int nlow, int hs, double rtol, double atol, bool init, double lo0,
                             double hi0, double* clo, double* chi, double* cval) {
  const int tid = threadIdx.x;
  const int lane = tid & (KPC_WAVE - 1);
  int G = ncand > 0 ? (int)blockDim.x / ncand : 1;
  G = G > 8 ? 8 : (G < 1 ? 1 : G);
  int Gp = 1;
  while (Gp * 2 <= G) Gp *= 2;  // power of two: groups align with lanes
  G = Gp;
  const int c = tid / G, g = tid % G;
  const bool own = c < ncand;
  const int idx = own ? (c < nlow ? c : hs + (c - nlow)) : 0;
  double lo = lo0, hi = hi0;
  if (own && !init) {
    lo = clo[c];
    hi = chi[c];
  }
  for (int round = 0; round < 400; ++round) {
    const double w = hi - lo;
    const bool done = !own || w <= rtol * fmax(fabs(lo), fabs(hi)) + atol + pivmin;
    if (__syncthreads_and(done ? 1 : 0)) break;
    const double step = w / (double)(G + 1);
    int flag = 0;
    if (own && !done) {
      const double xg = lo + step * (double)(g + 1);
      flag = lz_sturm(al, e2, m, xg, pivmin) > idx;
    }
    // first flagged point of the group (count(x) > idx moves the upper end)
    const unsigned long long bal = __ballot(flag);
    const int gbase = lane - g;
    const unsigned bits = (unsigned)((bal >> gbase) & ((1ULL << G) - 1ULL));
    if (own && !done) {
      const int first = bits ? __builtin_ctz(bits) : G;
      const double nlo = first == 0 ? lo : lo + step * (double)first;
      const double nhi = first == G ? hi : lo + step * (double)(first + 1);
      if (nhi > nlo) {
        lo = nlo;
        hi = nhi;
      } else {
        hi = lo;  // no representable interior point left
      }
    }
  }
  if (own && g == 0) {
    clo[c] = lo;
    chi[c] = hi;
    cval[c] = 0.5 * (lo + hi);
  }
  __syncthreads();
}

// Z[:, e] = Q S[:, e] for e < k on FP64 matrix cores (v_mfma_f64_16x16x4f64):
// a wave takes tiles of 16 rows and runs K over the basis four vectors per
// step; the A operand (lane l: row l % 16, basis vector l / 16 of the step)
// is loaded once per step and shared by the NTC tiles of 16 vectors (B: lane
// l holds coefficient b0 + l / 16 of vector l % 16).  Accumulator item v of
// lane l is row 4 v + l / 16, vector l % 16.  Rows past n load a valid row
// (not stored); basis vectors past m and vectors past k load 0.  Row i of the
// result goes to row perm[i] of Z (row i when perm is null).
typedef double lz_v4d __attribute__((ext_vector_type(4)));
template <int NTC>
__device__ inline void ritz_mfma(const double* Q, const double* Sv, int sst, int n, int m, int k,
                                 const int* perm, double* Z) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int wid = threadIdx.x / KPC_WAVE;
  const int nw = (int)blockDim.x / KPC_WAVE;
  const int r16 = lane & 15, kq = lane >> 4;
  const int ntr = (n + 15) >> 4;
  for (int tr = wid; tr < ntr; tr += nw) {
    const int row = tr * 16 + r16;
    const double* qr = Q + (row < n ? row : n - 1);
    lz_v4d acc[NTC];
#pragma unroll
    for (int t = 0; t < NTC; ++t) acc[t] = lz_v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int b0 = 0; b0 < m; b0 += 4) {
      const int b = b0 + kq;
      const double a = b < m ? qr[(size_t)b * n] : 0.0;
#pragma unroll
      for (int t = 0; t < NTC; ++t) {
        const int c = t * 16 + r16;
        const double bv = (b < m && c < k) ? Sv[(size_t)c * sst + b] : 0.0;
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[t], 0, 0, 0);
      }
    }
#pragma unroll
    for (int t = 0; t < NTC; ++t) {
      const int c = t * 16 + r16;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = tr * 16 + 4 * v + kq;
        if (i < n && c < k) Z[(size_t)c * n + (perm ? perm[i] : i)] = acc[t][v];
      }
    }
  }
}
__device__ inline void ritz_assemble_mfma(const double* Q, const double* Sv, int sst, int n, int m,
                                          int k, const int* perm, double* Z) {
  switch ((k + 15) >> 4) {
    case 1: ritz_mfma<1>(Q, Sv, sst, n, m, k, perm, Z); break;
    case 2: ritz_mfma<2>(Q, Sv, sst, n, m, k, perm, Z); break;
    case 3: ritz_mfma<3>(Q, Sv, sst, n, m, k, perm, Z); break;
    default: ritz_mfma<4>(Q, Sv, sst, n, m, k, perm, Z); break;
  }
}

}  // namespace kpc
