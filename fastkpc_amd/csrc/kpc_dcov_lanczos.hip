// dcov.gamma components by Lanczos on the implicit distance matrix.
//
// Same output contract as dcov_component_kernel (top-k |lambda| eigenpairs
// of D_ij = |r_i - r_j| selected with the higher-index tie break, centred
// vectors, sum(D) and the self moment; kpcalg/R/dcovgamma.R:67-79,
// fastkpc/src/legacy_dcov_gamma_cpp.cpp:60-78, :135, :283-300), but D is
// never stored: the residual is sorted once, and on sorted values
// (D v)_i = x_i (2 P_i - sum v) - 2 X_i + sum x v with P, X the exclusive
// prefix sums of v and x v, so a Lanczos step costs one block scan (O(n))
// instead of n^2 products.  Full reorthogonalisation (two
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

// Dynamic LDS bytes of one workgroup when two share a CU's 160 KB: 80 KB less
// 512 bytes for the kernel's static LDS (256 bytes).
constexpr int kLzLdsBudget = 81920 - 512;

// Doubles of the LDS workspace: the sort keys (2 P, P = n rounded up to a
// power of two, at most 4096), the twisted-factorisation pivots of a
// convergence check (2 m per eigenpair), the Ritz coefficients and the
// moment's row blocks, and between convergence checks the first basis
// columns.  It takes the rest of the budget, at least `base`.
__host__ __device__ inline int lanczos_ivw(int n, int mmax, int k) {
  (void)k;
  const int nl = (n + 63) & ~63;
  const int ml = (mmax + 64) & ~63;
  const int base = 4096 > 6 * mmax + 64 ? 4096 : 6 * mmax + 64;
  const int room =
      (kLzLdsBudget - (3 * kMaxNumCol + 8) * 4) / 8 - (4 * nl + 4 * ml + 32 + 8 * kMaxNumCol);
  return base > room ? base : room;
}

// Top-kk candidates by |value| (ties toward the higher ascending index): a
// candidate's rank is the number of candidates ahead of it in that order, and
// ranks below kk fill sval/sidx in order (one thread per candidate).
__device__ void lz_select(const double* cval, int ncand, int nlow, int hs, int kk, double* sval,
                          int* sidx) {
  for (int c = threadIdx.x; c < ncand; c += blockDim.x) {
    const double a = fabs(cval[c]);
    const int ic = c < nlow ? c : hs + (c - nlow);
    int rank = 0;
    for (int o = 0; o < ncand; ++o) {
      const double b = fabs(cval[o]);
      const int io = o < nlow ? o : hs + (o - nlow);
      rank += (b > a) || (b == a && io > ic);
    }
    if (rank < kk) {
      sval[rank] = cval[c];
      sidx[rank] = c;
    }
  }
  __syncthreads();
}

// Gram matrix G = Z' Z of the k vectors (columns of Z, n rows) on FP64
// matrix cores, into G (k x k, LDS): tiles of 16 x 16 with i <= j, one per
// wave in turn, K over the rows four per step; an entry's sum runs over rows
// 0 .. n - 1 as one chain of FMAs.  Vectors past k and rows past n load 0
// (vector indices clamped to a valid one).
template <int NTC>
__device__ inline void gram_mfma(const double* Z, int n, int k, double* G) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int wid = threadIdx.x / KPC_WAVE;
  const int nw = (int)blockDim.x / KPC_WAVE;
  const int r16 = lane & 15, kq = lane >> 4;
  int tile = 0;
  for (int ti = 0; ti < NTC; ++ti)
    for (int tj = ti; tj < NTC; ++tj, ++tile) {
      if (tile % nw != wid) continue;
      const int ci = ti * 16 + r16, cj = tj * 16 + r16;
      const double* zi = Z + (size_t)(ci < k ? ci : k - 1) * n;
      const double* zj = Z + (size_t)(cj < k ? cj : k - 1) * n;
      const bool oki = ci < k, okj = cj < k;
      lz_v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
      for (int r0 = 0; r0 < n; r0 += 4) {
        const int r = r0 + kq;
        const int rr = r < n ? r : n - 1;
        const double a = (r < n && oki) ? zi[rr] : 0.0;
        const double b = (r < n && okj) ? zj[rr] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gi = ti * 16 + 4 * v + kq, gj = tj * 16 + r16;
        if (gi < k && gj < k) {
          G[gi * k + gj] = acc[v];
          G[gj * k + gi] = acc[v];
        }
      }
    }
}
__device__ inline void gram_assemble_mfma(const double* Z, int n, int k, double* G) {
  switch ((k + 15) >> 4) {
    case 1: gram_mfma<1>(Z, n, k, G); break;
    case 2: gram_mfma<2>(Z, n, k, G); break;
    case 3: gram_mfma<3>(Z, n, k, G); break;
    default: gram_mfma<4>(Z, n, k, G); break;
  }
}

// Wave sum by DPP moves: quad xor 1, quad xor 2, half-row mirror and row
// mirror give every lane its 16-lane row total in registers, then the four
// row totals are read (lanes 0, 16, 32, 48) and added as (r0 + r1) + (r2 + r3),
// the same value in every lane.  All 64 lanes must be active.
template <int kCtrl>
__device__ __forceinline__ double lz_dpp(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), kCtrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), kCtrl, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lz_readlane(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += lz_dpp<0xB1>(v);   // quad_perm [1,0,3,2]
  v += lz_dpp<0x4E>(v);   // quad_perm [2,3,0,1]
  v += lz_dpp<0x141>(v);  // row_half_mirror
  v += lz_dpp<0x140>(v);  // row_mirror
  return (lz_readlane(v, 0) + lz_readlane(v, 16)) + (lz_readlane(v, 32) + lz_readlane(v, 48));
}

// Inclusive prefix sums over the wave of two values at once by DPP moves:
// row_shr 1, 2, 4, 8 within each 16-lane row (lanes past the row start read
// 0), then row_bcast 15 (row 0's total into row 1, row 2's into row 3) and
// row_bcast 31 (the total of rows 0-1 into rows 2 and 3).  All 64 lanes must
// be active.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ double lz_dpp_masked(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), kCtrl, kRowMask, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), kCtrl, kRowMask, 0xF, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ void wave_scan2_dpp(double& a, double& b) {
  a += lz_dpp_masked<0x111, 0xF>(a);  // row_shr:1
  b += lz_dpp_masked<0x111, 0xF>(b);
  a += lz_dpp_masked<0x112, 0xF>(a);  // row_shr:2
  b += lz_dpp_masked<0x112, 0xF>(b);
  a += lz_dpp_masked<0x114, 0xF>(a);  // row_shr:4
  b += lz_dpp_masked<0x114, 0xF>(b);
  a += lz_dpp_masked<0x118, 0xF>(a);  // row_shr:8
  b += lz_dpp_masked<0x118, 0xF>(b);
  a += lz_dpp_masked<0x142, 0xA>(a);  // row_bcast:15 into rows 1, 3
  b += lz_dpp_masked<0x142, 0xA>(b);
  a += lz_dpp_masked<0x143, 0xC>(a);  // row_bcast:31 into rows 2, 3
  b += lz_dpp_masked<0x143, 0xC>(b);
}

// block_sum with one barrier: the waves' partial sums go to one of two slot
// sets, alternating call by call (par), so a call never overwrites slots that a
// thread may still be reading: the call before it (other set) ends with a
// barrier that every thread reaches only after its reads of this set two
// calls back.  Same wave sums and the same order over the waves as
// block_sum.  All threads make the same sequence of calls.
__device__ __forceinline__ double block_sum_alt(double v, double* slots, int& par) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int wid = threadIdx.x / KPC_WAVE;
  const int nw = (blockDim.x + KPC_WAVE - 1) / KPC_WAVE;
  double* r = slots + (par ? 8 : 0);
  par ^= 1;
  v = wave_sum_dpp(v);
  if (lane == 0) r[wid] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < nw; ++w) t += r[w];
  return t;
}

// Coefficients of z against basis columns b0 .. b0 + nb - 1 (lane partial
// sums over rows lane, lane + 64, ...; columns past nb repeat the last one)
// with the columns read from qbase, the slab or its LDS copy.
__device__ __forceinline__ void gs_block_coeffs(const double* qbase, const double* z, int n,
                                                int b0, int nb, int lane, double g[8]) {
  const double* qrow[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) qrow[c] = qbase + (size_t)(b0 + (c < nb ? c : nb - 1)) * n;
#pragma unroll 2
  for (int i = lane; i < n; i += KPC_WAVE) {
    double v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = qrow[c][i];
    const double zi = z[i];
#pragma unroll
    for (int c = 0; c < 8; ++c) g[c] += zi * v[c];
  }
}

// Row i's combination sum_b cg[b] q_b[i] over whole blocks b0 <= b < b1 of
// eight columns read from qbase, accumulated into sa[b mod 8].
__device__ __forceinline__ void gs_block_sums(const double* qbase, const double* cg, int n, int i,
                                              int b0, int b1, double sa[8]) {
#pragma unroll 2
  for (int b = b0; b + 8 <= b1; b += 8) {
    double v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = qbase[(size_t)(b + c) * n + i];
#pragma unroll
    for (int c = 0; c < 8; ++c) sa[c] += cg[b + c] * v[c];
  }
}

__global__ __launch_bounds__(kLzThreads, 4) void dcov_lanczos_kernel(
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
  double* rsum = red + 16;             // 2 x 8: alternating block-sum slots of the steps
  double* cval = rsum + 16;            // 2 kMaxNumCol
  double* sval = cval + 2 * kMaxNumCol;   // kMaxNumCol
  double* sres = sval + kMaxNumCol;       // kMaxNumCol
  int* cidx = reinterpret_cast<int*>(sres + kMaxNumCol);  // 2 kMaxNumCol
  int* sidx = cidx + 2 * kMaxNumCol;                      // kMaxNumCol
  int* ictl = sidx + kMaxNumCol;                          // 8
  double* clo = reinterpret_cast<double*>(ictl + 8);      // 2 kMaxNumCol brackets
  double* chi = clo + 2 * kMaxNumCol;                     // 2 kMaxNumCol
  double* ivl = chi + 2 * kMaxNumCol;                     // inverse-iteration workspace
  const int ivw = lanczos_ivw(n, mmax, k);
  int* perm = reinterpret_cast<int*>(ivl + ivw);          // n: sorted position -> row
  const int tid = threadIdx.x;
  const int lane = tid & (KPC_WAVE - 1);
  const int wid = tid / KPC_WAVE;
  const int nw = blockDim.x / KPC_WAVE;
  double* Q = scratch + (size_t)blockIdx.x * scratch_stride;   // n x (mmax + 1)
  double* S = Q + (size_t)n * (mmax + 1);                      // k x mmax (Ritz coefficients)
  // Between convergence checks the workspace past the scan slots also holds
  // basis columns 0 .. ncache0 - 1 (whole Gram-Schmidt blocks of eight), which
  // both Gram-Schmidt passes then read from LDS instead of the slab: the same
  // values, so the same results.  A convergence check overwrites them.
  double* qc = ivl + 64;
  int ncache0 = (ivw - 64) / n;
  if (ncache0 > mmax + 1) ncache0 = mmax + 1;
  ncache0 &= ~7;
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
    // Sort the residual once (bitonic, ties by row index): every Lanczos
    // vector lives in ascending-residual order, where D v is a prefix-sum
    // expression, and the Ritz vectors are permuted back on output.  The
    // sorted values are shifted by their median so the prefix form works on
    // centred values (D is shift invariant).
    {
      int P = 1;
      while (P < n) P <<= 1;
      double* key = ivl;      // P keys, then P row indices (ivw >= 2 kMaxN)
      double* kid = ivl + P;
      for (int i = tid; i < P; i += blockDim.x) {
        key[i] = i < n ? rq[i].x : __builtin_inf();
        kid[i] = (double)i;
      }
      __syncthreads();
      for (int kk = 2; kk <= P; kk <<= 1)
        for (int jj = kk >> 1; jj > 0; jj >>= 1) {
          for (int t = tid; t < P / 2; t += blockDim.x) {
            const int i = 2 * jj * (t / jj) + (t % jj), l = i + jj;
            const double a = key[i], b = key[l], ia = kid[i], ib = kid[l];
            const bool gt = a > b || (a == b && ia > ib);
            if (((i & kk) == 0) == gt) {
              key[i] = b;
              key[l] = a;
              kid[i] = ib;
              kid[l] = ia;
            }
          }
          __syncthreads();
        }
      const double mid = key[n / 2];
      for (int i = tid; i < n; i += blockDim.x) {
        perm[i] = (int)kid[i];
        rq[i].x = key[i] - mid;
      }
      __syncthreads();
    }
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
    int ncache = ncache0;
    for (int i = tid; i < n; i += blockDim.x) {
      const double v = z[perm[i]] * inv0;
      rq[i].y = v;
      Q[i] = v;
      if (ncache > 0) qc[i] = v;
    }
    __syncthreads();
    int m_final = mmax;
    int par = 0;  // slot set of the next block_sum_alt
    // first convergence check at 2k + 6 steps: top-k |lambda| components of
    // distance matrices converge there for n from 200 to 2000 (one check per
    // component); later checks follow the observed residual decay
    int next_check = 2 * k + 6 < mmax ? 2 * k + 6 : mmax, prev_m = 0;
    // n <= blockDim: thread t owns row t in every per-row loop of a step (the
    // matvec's runs are single rows), so the barriers after the matvec and
    // after the normalisation guard nothing: alpha's partial is formed in the
    // matvec loop (the same product), and every cross-row read is behind the
    // block sums' barriers
    const bool own_rows = n <= (int)blockDim.x;
    double prev_worst = 0.0;
    double tmax = 0.0;  // running max of |alpha_j| + |beta_j| (scale of T)
    if (tid == 0) { ictl[1] = 0; ictl[2] = 2; }
    __syncthreads();
    for (int j = 0; j < mmax; ++j) {
      double apart = 0.0;  // alpha's partial sum
      // z = D q_j on the sorted residual: with exclusive prefix sums
      // P_i = sum_{l<i} q_l and X_i = sum_{l<i} x_l q_l and totals Tq, Tx,
      // (D q)_i = x_i (2 P_i - Tq) - 2 X_i + Tx.  A thread scans a run of
      // consecutive rows; waves scan by lane shuffles, the wave totals meet
      // in LDS.
      {
        const int per = (n + (int)blockDim.x - 1) / (int)blockDim.x;
        const int i0 = min(tid * per, n), i1 = min(i0 + per, n);
        double sv = 0.0, sx = 0.0;
        for (int i = i0; i < i1; ++i) {
          const double2 e = rq[i];
          sv += e.y;
          sx += e.x * e.y;
        }
        double iv = sv, ix = sx;
        wave_scan2_dpp(iv, ix);
        double ev = __shfl_up(iv, 1, KPC_WAVE), ex = __shfl_up(ix, 1, KPC_WAVE);
        if (lane == 0) ev = ex = 0.0;
        double* scan = ivl;  // free between convergence checks
        if (lane == KPC_WAVE - 1) {
          scan[wid] = iv;
          scan[nw + wid] = ix;
        }
        __syncthreads();
        double bv = 0.0, bx = 0.0, tq = 0.0, tx = 0.0;
        for (int w = 0; w < nw; ++w) {
          const double a = scan[w], b = scan[nw + w];
          if (w < wid) {
            bv += a;
            bx += b;
          }
          tq += a;
          tx += b;
        }
        double pv = bv + ev, px = bx + ex;
        for (int i = i0; i < i1; ++i) {
          const double2 e = rq[i];
          const double zi = e.x * (2.0 * pv - tq) - 2.0 * px + tx;
          z[i] = zi;
          if (own_rows) apart += e.y * zi;
          pv += e.y;
          px += e.x * e.y;
        }
      }
      if (!own_rows) {
        __syncthreads();
        for (int i = tid; i < n; i += blockDim.x) apart += rq[i].y * z[i];
      }
      const double aj = block_sum_alt(apart, rsum, par);
      double part;
      // three-term recurrence
      const double bprev = j > 0 ? be[j - 1] : 0.0;
      const double* qp = Q + (size_t)(j > 0 ? j - 1 : 0) * n;
      part = 0.0;
      for (int i = tid; i < n; i += blockDim.x) {
        const double v = z[i] - (aj * rq[i].y + bprev * qp[i]);
        z[i] = v;
        part += v * v;
      }
      double zn2 = block_sum_alt(part, rsum, par);
      // Classical Gram-Schmidt against Q[0..j], repeated once more when a pass
      // leaves less than half the squared norm it started from (DGKS).  A
      // wave forms eight coefficients at once (lanes over rows, coalesced;
      // basis rows past j are clamped to j so every load is unconditional),
      // then each row subtracts its combination with eight accumulators.
      for (int pass = 0; pass < 2; ++pass) {
        for (int b0 = wid * 8; b0 <= j; b0 += nw * 8) {
          double g[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
          const int nb = min(8, j + 1 - b0);
          if (b0 < ncache)
            gs_block_coeffs(qc, z, n, b0, nb, lane, g);
          else
            gs_block_coeffs(Q, z, n, b0, nb, lane, g);
#pragma unroll
          for (int c = 0; c < 8; ++c) g[c] = wave_sum_dpp(g[c]);
          if (lane < nb) {
            double gv = g[0];
#pragma unroll
            for (int c = 1; c < 8; ++c) gv = lane == c ? g[c] : gv;
            cg[b0 + lane] = gv;
          }
        }
        __syncthreads();
        part = 0.0;
        for (int i = tid; i < n; i += blockDim.x) {
          double sa[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
          const int bc = min((j + 1) & ~7, ncache);  // whole blocks held in LDS
          gs_block_sums(qc, cg, n, i, 0, bc, sa);
          int b = bc;
#pragma unroll 2
          for (; b + 8 <= j + 1; b += 8) {
            double v[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) v[c] = Q[(size_t)(b + c) * n + i];
#pragma unroll
            for (int c = 0; c < 8; ++c) sa[c] += cg[b + c] * v[c];
          }
          for (; b <= j; ++b) sa[0] += cg[b] * (b < ncache ? qc[(size_t)b * n + i] : Q[(size_t)b * n + i]);
          const double zi =
              z[i] - (((sa[0] + sa[1]) + (sa[2] + sa[3])) + ((sa[4] + sa[5]) + (sa[6] + sa[7])));
          z[i] = zi;
          part += zi * zi;
        }
        const double after = block_sum_alt(part, rsum, par);
        const bool again = after < 0.5 * zn2;
        zn2 = after;
        if (!again) break;
      }
      if (tid == 0) al[j] = aj;
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
        if (j + 1 < ncache) qc[(size_t)(j + 1) * n + i] = v;
      }
      const int m = j + 1;
      const bool check = m >= next_check || m == mmax || breakdown;
      if (!own_rows || check) __syncthreads();
      if (!check) continue;
      // Gershgorin bounds of T
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
      // Ritz values to full precision: the kk lowest and the hcnt highest
      // eigenvalues of T.  When some computed high candidate is left out of
      // the top kk by |value|, every uncomputed middle eigenvalue is too (it
      // lies between the lowest-kk and the highest-hcnt sets, so either the
      // unselected high candidate or all kk low ones dominate it); otherwise
      // hcnt doubles.  A distance matrix has one positive eigenvalue, so the
      // first pass usually decides.
      int hcnt = kk < 4 ? kk : 4;
      int hs = 0;
      for (;;) {
        hs = (m - hcnt) > nlow ? (m - hcnt) : nlow;
        const int ncand = nlow + (m - hs);
        lz_multisect(al, e2, m, pivmin, ncand, nlow, hs, 2.0 * 2.220446049250313e-16,
                     2.0 * 2.220446049250313e-16 * tn, true, lo0, hi0, clo, chi, cval);
        lz_select(cval, ncand, nlow, hs, kk, sval, sidx);
        int high = 0;
        for (int c = 0; c < kk; ++c) high += sidx[c] >= nlow;
        if (high < m - hs || hs == nlow || hcnt >= kk) break;
        hcnt = 2 * hcnt < kk ? 2 * hcnt : kk;
        __syncthreads();
      }
      // eigenvectors of T by twisted factorisations (pivot workspace in LDS,
      // batches of eigenpairs), exact Ritz residuals |beta_m s_m|; the
      // coefficient vectors go to S at every check, so the last check's are
      // the final ones
      {
        const int per = 2 * m;
        const int B = ivw / per > 0 ? ivw / per : 1;
        for (int c0 = 0; c0 < kk; c0 += B) {
          const int c = c0 + tid;
          if (tid < B && c < kk) {
            double* wsl = ivl + (size_t)tid * per;
            const double last =
                lz_twisted(al, be, m, sval[c], pivmin, wsl, wsl + m, S + (size_t)c * mmax);
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
      if (ictl[0]) { m_final = m; break; }
      ncache = 0;  // the check's pivots overwrote the cached columns
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
    // Ritz vectors U = Q S -> output (coefficients staged in LDS when they fit)
    double* Z = comp_vecs + (size_t)slot * n * k;
    if ((long long)k * m_final <= ivw) {
      for (int e = tid; e < k * m_final; e += blockDim.x)
        ivl[e] = S[(size_t)(e / m_final) * mmax + e % m_final];
      __syncthreads();
      ritz_assemble_mfma(Q, ivl, m_final, n, m_final, k, perm, Z);
    } else {
      ritz_assemble_mfma(Q, S, mmax, n, m_final, k, perm, Z);
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
    // self moment sum_ij lambda_i lambda_j (z_i . z_j)^2 over the centred
    // vectors: the Gram entries (rows in order) on matrix cores into the
    // workspace, then one thread per pair (i <= j)
    gram_assemble_mfma(Z, n, k, ivl);
    __syncthreads();
    const int npairs = k * (k + 1) / 2;
    double mpart = 0.0;
    for (int pidx0 = tid; pidx0 < npairs; pidx0 += blockDim.x) {
      int pidx = pidx0, i = 0;
      while (pidx >= k - i) {
        pidx -= k - i;
        ++i;
      }
      const int jj = i + pidx;
      const double gv = ivl[i * k + jj];
      mpart += (jj == i ? 1.0 : 2.0) * (sval[i] * sval[jj] * gv * gv);
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

size_t dcov_lanczos_lds_bytes(int n, int k, int mmax) {
  const int nl = (n + 63) & ~63;
  const int ml = (mmax + 64) & ~63;
  return (size_t)(4 * nl + 4 * ml + 32 + 4 * kMaxNumCol + 4 * kMaxNumCol + lanczos_ivw(n, mmax, k)) *
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
  if (n < 1 || n > kMaxN || k < 1 || k > kMaxNumCol) return hipErrorInvalidValue;
  const size_t lds = dcov_lanczos_lds_bytes(n, k, mmax);
  hipError_t err = hipFuncSetAttribute((const void*)dcov_lanczos_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(dcov_lanczos_kernel, dim3(grid), dim3(kLzThreads), lds, stream, src_base,
                     src_off, dst_slot, count, n, k, vals, vecs, dsum, moment, status, scratch,
                     dcov_lanczos_scratch_doubles(n, k, mmax), mmax, tol);
  return hipGetLastError();
}

}  // namespace kpc
