// hsic.gamma, hsic.perm and dcc.perm for gfx950.
//
// Components (one per unique (target, S) residual, cached by the host):
//   hsic.*   kernlab-inchol factor of the Gaussian kernel matrix
//            (kpcalg/R/hsicgamma.R:60-66; fastkpc/src/cuda/full_cuda_ci_method_batch.cu:
//            445-658): sigma = 1 / sig, tol 0.001, rank <= numCol, first pivot
//            row 0 then the first maximum of the diagonal residues; stored
//            column-centred with the off-diagonal kernel sum (uncentred) and
//            the self moment |Gc^T Gc|_F^2.  One 256-lane workgroup per
//            component; every lane owns up to 8 rows whose residues stay in
//            registers.  The pivot row of the factor, G[p, 0..rank), is the
//            forward solve of the reference (mathematically identical; the
//            factor rows are already computed), so no serial triangular solve
//            sits on the critical path.
//   dcc.perm residual plus the row means of |r_i - r_j|^index and the grand mean
//            (fastkpc/src/cuda/full_cuda_ci_method_batch.cu:284-339); the
//            double-centred matrix is never materialised.
//
// Pairs (one workgroup per test, strided):
//   hsic     HSIC = |Gcx^T Gcy|_F^2 / n^2; gamma moments and tail
//            (kpcalg/R/hsicgamma.R:74-83), or R permutation replicates of the
//            second factor's rows (kpcalg/R/hsicperm.R:77-88).
//   dcc.perm T = sum_ij A_ij B_ij and replicates sum_ij A_ij B_{pi(i) pi(j)},
//            A and B double-centred distances evaluated on the fly from LDS
//            over the upper triangle of 64 x 64 tiles (off-diagonal tiles
//            count twice); statistic T / n, p = exceed / (R + include_observed)
//            (fastkpc/src/cuda/full_cuda_ci_method_batch.cu:341-443).
//
// Permutations are generated in LDS by the workgroup itself, one replicate per
// lane: Durstenfeld shuffle driven by Philox4x32-10 (key = task seed, counter
// {i >> 1, replicate, 0x4b504331, 0}), the same index arrays the oracle draws.
#include "kpc_common.h"
#include "kpc_device.h"
#include "kpc_gamma.h"

namespace kpc {

constexpr int kHsicThreads = 256;
constexpr int kRowsPerThread = kMaxN / kHsicThreads;
constexpr double kIncholTol = 0.001;
constexpr int kPermLdsBytes = 48 * 1024;

// ------------------------------------------------------------ permutations
__device__ __forceinline__ void philox4x32_10(uint32_t& c0, uint32_t& c1, uint32_t& c2,
                                              uint32_t& c3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
    const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
}

template <typename T>
__device__ void make_permutation(unsigned long long seed, int rep, int n, T* perm) {
  for (int i = 0; i < n; ++i) perm[i] = (T)i;
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
  int blk = -1;
  for (int i = n - 1; i >= 1; --i) {
    if ((i >> 1) != blk) {
      blk = i >> 1;
      w0 = (uint32_t)blk;
      w1 = (uint32_t)rep;
      w2 = 0x4b504331u;
      w3 = 0u;
      philox4x32_10(w0, w1, w2, w3, k0, k1);
    }
    const unsigned long long u = (i & 1) ? ((unsigned long long)w3 << 32 | w2)
                                         : ((unsigned long long)w1 << 32 | w0);
    const int j = (int)__umul64hi(u, (unsigned long long)(i + 1));
    const T t = perm[i];
    perm[i] = perm[j];
    perm[j] = t;
  }
}

__global__ void permutation_table_kernel(const unsigned long long* __restrict__ seeds, int count,
                                         int n, int replicates, int* __restrict__ out) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long long)count * replicates) return;
  const int item = (int)(g / replicates), rep = (int)(g % replicates);
  make_permutation<int>(seeds[item], rep, n, out + g * n);
}

// First maximum over the workgroup: largest value, smallest index on ties.
__device__ __forceinline__ void block_argmax_first(double& v, int& idx, double* rv, int* ri) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int wid = threadIdx.x / KPC_WAVE;
  const int nw = blockDim.x / KPC_WAVE;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(v, off, KPC_WAVE);
    const int oi = __shfl_xor(idx, off, KPC_WAVE);
    if (ov > v || (ov == v && oi < idx)) {
      v = ov;
      idx = oi;
    }
  }
  __syncthreads();
  if (lane == 0) {
    rv[wid] = v;
    ri[wid] = idx;
  }
  __syncthreads();
  v = rv[0];
  idx = ri[0];
  for (int w = 1; w < nw; ++w)
    if (rv[w] > v || (rv[w] == v && ri[w] < idx)) {
      v = rv[w];
      idx = ri[w];
    }
}

// ------------------------------------------------------------------ inchol
__global__ __launch_bounds__(kHsicThreads) void inchol_kernel(
    const double* __restrict__ src_base, const long long* __restrict__ src_off,
    const int* __restrict__ dst_slot, int count, int n, int k, double sigma,
    double* __restrict__ vecs, double* __restrict__ offdiag, double* __restrict__ moment,
    int* __restrict__ rank_out, double* __restrict__ margin_out) {
  __shared__ double s_row[kMaxNumCol];
  __shared__ double red[16];
  __shared__ int redi[16];
  __shared__ double colacc[2 * kHsicThreads / KPC_WAVE];
  const int tid = threadIdx.x;
  const int lane = tid & (KPC_WAVE - 1);
  const int wid = tid / KPC_WAVE;
  const int nw = blockDim.x / KPC_WAVE;
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const double* x = src_base + src_off[item];
    const int slot = dst_slot[item];
    double* G = vecs + (size_t)slot * n * k;
    double xr[kRowsPerThread], dg[kRowsPerThread];
#pragma unroll
    for (int r = 0; r < kRowsPerThread; ++r) {
      const int i = tid + r * kHsicThreads;
      xr[r] = i < n ? x[i] : 0.0;
      dg[r] = 1.0;
    }
    int rank = 0, piv = 0;
    double residue = 1.0, mg = 1e300;
    while (residue > kIncholTol && rank < k) {
      const double xp = x[piv];
      double t2 = 1.0;  // k(p, p) = exp(-0)
      for (int c = 0; c < rank; ++c) t2 -= s_row[c] * s_row[c];
      if (!(t2 > 0.0) || !isfinite(t2)) break;
      const double tau = sqrt(t2);
      double lbest = -INFINITY;
      int lidx = 0x7fffffff;
#pragma unroll
      for (int r = 0; r < kRowsPerThread; ++r) {
        const int i = tid + r * kHsicThreads;
        if (i < n) {
          const double d = xr[r] - xp;
          const double col = exp(-sigma * d * d);
          double proj = 0.0;
          for (int c = 0; c < rank; ++c) proj += G[i + (size_t)n * c] * s_row[c];
          const double u = (col - proj) / tau;
          G[i + (size_t)n * rank] = u;
          dg[r] -= u * u;
          if (dg[r] > lbest) {
            lbest = dg[r];
            lidx = i;
          }
        }
      }
      ++rank;
      double best = lbest;
      int bi = lidx;
      block_argmax_first(best, bi, red, redi);
      double lsec = -INFINITY;
#pragma unroll
      for (int r = 0; r < kRowsPerThread; ++r) {
        const int i = tid + r * kHsicThreads;
        if (i < n && dg[r] < best && dg[r] > lsec) lsec = dg[r];
      }
      __syncthreads();
      const double second = block_max(lsec, red + 8);
      // non-finite residues (a failed fit) end the factorisation here; the
      // host rejects such a fit before any statistic is used
      const bool valid = bi >= 0 && bi < n && !isnan(best);
      residue = valid ? best : -INFINITY;
      piv = valid ? bi : 0;
      // decision margins relative to the mass rounding can move (a residue is
      // 1 - sum g^2): stop rule residue > tol, and the pivot choice
      if (rank < k) {
        mg = fmin(mg, fabs(residue - kIncholTol) / fmax(1.0 - residue, kIncholTol));
        if (residue > kIncholTol) {
          const double mass = fmax((1.0 - best) + (1.0 - second), 2.220446049250313e-16);
          mg = fmin(mg, (best - second) / mass);
        }
      }
      if (residue > kIncholTol && rank < k) {
        for (int c = tid; c < rank; c += blockDim.x) s_row[c] = G[piv + (size_t)n * c];
        __syncthreads();
      }
    }
    // zero the unused columns, then off-diagonal sum and centring per column
    for (int c = rank; c < k; ++c)
      for (int i = tid; i < n; i += blockDim.x) G[i + (size_t)n * c] = 0.0;
    double wtot = 0.0, wsq = 0.0;
    __syncthreads();
    for (int c = wid; c < rank; c += nw) {
      double* g = G + (size_t)n * c;
      double cs = 0.0, sq = 0.0;
      for (int i = lane; i < n; i += KPC_WAVE) {
        const double v = g[i];
        cs += v;
        sq += v * v;
      }
      cs = wave_sum(cs);
      sq = wave_sum(sq);
      wtot += cs * cs;
      wsq += sq;
      const double mean = cs / (double)n;
      for (int i = lane; i < n; i += KPC_WAVE) g[i] -= mean;
    }
    if (lane == 0) {
      colacc[2 * wid] = wtot;
      colacc[2 * wid + 1] = wsq;
    }
    __syncthreads();
    // self moment sum_ab (Gc^T Gc)_ab^2 over a <= b (off-diagonal twice)
    double wm = 0.0;
    const int npair = rank * (rank + 1) / 2;
    for (int e = wid; e < npair; e += nw) {
      int a = 0, rem = e;
      while (rem >= rank - a) {
        rem -= rank - a;
        ++a;
      }
      const int b = a + rem;
      const double* ga = G + (size_t)n * a;
      const double* gb = G + (size_t)n * b;
      double s = 0.0;
      for (int i = lane; i < n; i += KPC_WAVE) s += ga[i] * gb[i];
      s = wave_sum(s);
      wm += (a == b ? 1.0 : 2.0) * s * s;
    }
    const double m = block_sum(lane == 0 ? wm : 0.0, red);
    if (tid == 0) {
      double tot = 0.0, sq = 0.0;
      for (int w = 0; w < nw; ++w) {
        tot += colacc[2 * w];
        sq += colacc[2 * w + 1];
      }
      offdiag[slot] = tot - sq;
      moment[slot] = m;
      rank_out[slot] = rank;
      margin_out[slot] = mg;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------- hsic pairs
// Sum over this thread's (a, b) entries of (sum_i A[i,a] B[perm(i),b])^2.
template <bool kPerm>
__device__ __forceinline__ double hsic_cross_part(const double* __restrict__ A, int ra,
                                                  const double* __restrict__ B, int rb, int n,
                                                  const unsigned short* perm) {
  double part = 0.0;
  const int ne = ra * rb;
  for (int e = threadIdx.x; e < ne; e += blockDim.x) {
    const int a = e / rb, b = e - (e / rb) * rb;
    const double* ga = A + (size_t)n * a;
    const double* gb = B + (size_t)n * b;
    double dot = 0.0;
    for (int i = 0; i < n; ++i) dot += ga[i] * gb[kPerm ? (int)perm[i] : i];
    part += dot * dot;
  }
  return part;
}

struct PermTally {
  int exceed;
  double mean, m2, margin;
};

__device__ __forceinline__ void tally(PermTally& t, double v, double observed, double scale,
                                      int count_before) {
  if (v >= observed) ++t.exceed;
  const double g = fabs(v - observed) / (fabs(observed) > 0.0 ? fabs(observed) : 1.0);
  t.margin = fmin(t.margin, g);
  const double s = v * scale;
  const double d = s - t.mean;
  t.mean += d / (double)(count_before + 1);
  t.m2 += d * (s - t.mean);
}

__device__ __forceinline__ void perm_record(double* o, const PermTally& t, double observed,
                                            double scale, int replicates, int include_observed,
                                            bool one_if_nonpositive) {
  const int denom = replicates + (include_observed ? 1 : 0);
  double p = denom > 0 ? (double)t.exceed / (double)denom : 1.0;
  if (one_if_nonpositive && !(observed > 0.0)) p = 1.0;
  o[0] = p;
  o[1] = observed * scale;
  o[2] = t.mean;
  o[3] = replicates > 1 ? t.m2 / (double)(replicates - 1) : 0.0;
  o[4] = t.margin;
}

__global__ __launch_bounds__(kHsicThreads) void hsic_pair_kernel(
    const int* __restrict__ pa, const int* __restrict__ pb,
    const unsigned long long* __restrict__ seeds, int count, int n, int k,
    const double* __restrict__ vecs, const int* __restrict__ rank,
    const double* __restrict__ offdiag, const double* __restrict__ moment, int permute,
    int replicates, int include_observed, int rbatch, double* __restrict__ out) {
  extern __shared__ unsigned short perm_lds[];
  __shared__ double red[16];
  const int tid = threadIdx.x;
  const double nd = (double)n;
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const int ca = pa[item], cb = pb[item];
    const int ra = rank[ca], rb = rank[cb];
    const double* A = vecs + (size_t)ca * n * k;
    const double* B = vecs + (size_t)cb * n * k;
    const double observed = block_sum(hsic_cross_part<false>(A, ra, B, rb, n, nullptr), red) /
                            (nd * nd);
    double* o = out + (size_t)item * kPairOut;
    if (!permute) {
      if (tid == 0) {
        const double mux = offdiag[ca] / (nd * (nd - 1.0));
        const double muy = offdiag[cb] / (nd * (nd - 1.0));
        const double mean = (1.0 + mux * muy - mux - muy) / nd;
        const double f =
            2.0 * (nd - 4.0) * (nd - 5.0) / (nd * (nd - 1.0) * (nd - 2.0) * (nd - 3.0));
        const double var = f * moment[ca] * moment[cb] / (nd * nd * nd * nd);
        const bool ok = isfinite(observed) && isfinite(mean) && isfinite(var) && mean > 0.0 &&
                        var > 0.0;
        o[0] = ok ? gamma_pvalue(observed, mean, var) : 1.0;
        o[1] = observed;
        o[2] = mean;
        o[3] = var;
        o[4] = 1e300;
      }
      __syncthreads();
      continue;
    }
    PermTally t{include_observed ? 1 : 0, 0.0, 0.0, 1e300};
    const unsigned long long seed = seeds[item];
    for (int base = 0; base < replicates; base += rbatch) {
      const int nb = min(rbatch, replicates - base);
      __syncthreads();
      if (tid < nb) make_permutation<unsigned short>(seed, base + tid, n, perm_lds + (size_t)tid * n);
      __syncthreads();
      for (int r = 0; r < nb; ++r) {
        const double v =
            block_sum(hsic_cross_part<true>(A, ra, B, rb, n, perm_lds + (size_t)r * n), red) /
            (nd * nd);
        if (tid == 0) tally(t, v, observed, 1.0, base + r);
      }
    }
    if (tid == 0) perm_record(o, t, observed, 1.0, replicates, include_observed, false);
    __syncthreads();
  }
}

// ---------------------------------------------------------------- dcc.perm
template <bool kUnitIndex>
__device__ __forceinline__ double dist_pow(double d, double index) {
  return kUnitIndex ? fabs(d) : pow(fabs(d), index);
}

template <bool kUnitIndex>
__global__ __launch_bounds__(kHsicThreads) void dcc_rowmean_kernel(
    const double* __restrict__ src_base, const long long* __restrict__ src_off,
    const int* __restrict__ dst_slot, int count, int n, int stride, double index,
    double* __restrict__ vecs, double* __restrict__ grand) {
  extern __shared__ double xs[];
  __shared__ double red[16];
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const double* x = src_base + src_off[item];
    double* V = vecs + (size_t)dst_slot[item] * stride;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      xs[i] = x[i];
      V[i] = x[i];
    }
    __syncthreads();
    double part = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const double xi = xs[i];
      double r = 0.0;
      for (int j = 0; j < n; ++j) r += dist_pow<kUnitIndex>(xi - xs[j], index);
      V[n + i] = r / (double)n;
      part += r;
    }
    const double total = block_sum(part, red);
    if (threadIdx.x == 0) grand[dst_slot[item]] = total / ((double)n * (double)n);
    __syncthreads();
  }
}

// sum_ij A_ij B_ij over the upper triangle of 64 x 64 tiles, A_ij =
// ((D_ij - a_i) - a_j) + abar from XA = {x, a} and likewise B from YB.
template <bool kUnitIndex>
__device__ __forceinline__ double dcc_tile_sum(const double2* XA, const double2* YB, int n,
                                               double abar, double bbar, double index) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int wid = threadIdx.x / KPC_WAVE;
  const int nw = blockDim.x / KPC_WAVE;
  const int ns = (n + KPC_WAVE - 1) / KPC_WAVE;
  const int npair = ns * (ns + 1) / 2;
  double acc = 0.0;
  for (int sp = wid; sp < npair; sp += nw) {
    int s = 0, rem = sp;
    while (rem >= ns - s) {
      rem -= ns - s;
      ++s;
    }
    const int s2 = s + rem;
    const int i = s * KPC_WAVE + lane;
    double tile = 0.0;
    if (i < n) {
      const double2 xa = XA[i];
      const double2 yb = YB[i];
      const int j0 = s2 * KPC_WAVE;
      const int j1 = min(n, j0 + KPC_WAVE);
      for (int j = j0; j < j1; ++j) {
        const double2 xj = XA[j];
        const double2 yj = YB[j];
        const double a = ((dist_pow<kUnitIndex>(xa.x - xj.x, index) - xa.y) - xj.y) + abar;
        const double b = ((dist_pow<kUnitIndex>(yb.x - yj.x, index) - yb.y) - yj.y) + bbar;
        tile += a * b;
      }
    }
    acc += s == s2 ? tile : 2.0 * tile;
  }
  return acc;
}

template <bool kUnitIndex>
__global__ __launch_bounds__(kHsicThreads) void dcc_perm_kernel(
    const int* __restrict__ pa, const int* __restrict__ pb,
    const unsigned long long* __restrict__ seeds, int count, int n, int stride, double index,
    const double* __restrict__ vecs, const double* __restrict__ grand, int replicates,
    int include_observed, int rbatch, double* __restrict__ out) {
  extern __shared__ double2 dlds[];
  double2* XA = dlds;
  double2* YB = XA + n;
  double2* YP = YB + n;
  unsigned short* perm = (unsigned short*)(YP + n);
  __shared__ double red[16];
  const int tid = threadIdx.x;
  const double inv_n = 1.0 / (double)n;
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const int ca = pa[item], cb = pb[item];
    const double* VA = vecs + (size_t)ca * stride;
    const double* VB = vecs + (size_t)cb * stride;
    const double abar = grand[ca], bbar = grand[cb];
    for (int i = tid; i < n; i += blockDim.x) {
      XA[i] = make_double2(VA[i], VA[n + i]);
      YB[i] = make_double2(VB[i], VB[n + i]);
    }
    __syncthreads();
    const double observed =
        block_sum(dcc_tile_sum<kUnitIndex>(XA, YB, n, abar, bbar, index), red);
    PermTally t{include_observed ? 1 : 0, 0.0, 0.0, 1e300};
    const unsigned long long seed = seeds[item];
    for (int base = 0; base < replicates; base += rbatch) {
      const int nb = min(rbatch, replicates - base);
      __syncthreads();
      if (tid < nb) make_permutation<unsigned short>(seed, base + tid, n, perm + (size_t)tid * n);
      __syncthreads();
      for (int r = 0; r < nb; ++r) {
        const unsigned short* pr = perm + (size_t)r * n;
        for (int i = tid; i < n; i += blockDim.x) YP[i] = YB[pr[i]];
        __syncthreads();
        const double v = block_sum(dcc_tile_sum<kUnitIndex>(XA, YP, n, abar, bbar, index), red);
        if (tid == 0) tally(t, v, observed, inv_n, base + r);
      }
    }
    if (tid == 0)
      perm_record(out + (size_t)item * kPairOut, t, observed, inv_n, replicates, include_observed,
                  true);
    __syncthreads();
  }
}

// ---------------------------------------------------------------- launchers
static int perm_batch(int n, int replicates) {
  int rb = kPermLdsBytes / (2 * n);
  rb = rb < 1 ? 1 : rb;
  rb = rb > kHsicThreads ? kHsicThreads : rb;
  return replicates < rb ? (replicates > 0 ? replicates : 1) : rb;
}

hipError_t launch_inchol_components(const double* src_base, const long long* src_off,
                                    const int* dst_slot, int count, int n, int k, double sigma,
                                    double* vecs, double* offdiag, double* moment, int* rank,
                                    double* margin, int grid, hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  if (n > kMaxN || k > kMaxNumCol) return hipErrorInvalidValue;
  hipLaunchKernelGGL(inchol_kernel, dim3(grid), dim3(kHsicThreads), 0, stream, src_base, src_off,
                     dst_slot, count, n, k, sigma, vecs, offdiag, moment, rank, margin);
  return hipGetLastError();
}

hipError_t launch_hsic_pairs(const int* pa, const int* pb, const unsigned long long* seeds,
                             int count, int n, int k, const double* vecs, const int* rank,
                             const double* offdiag, const double* moment, int permute,
                             int replicates, int include_observed, double* out, int grid,
                             hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  const int rb = perm_batch(n, replicates);
  const size_t lds = permute && replicates > 0 ? (size_t)rb * n * sizeof(unsigned short) : 0;
  hipLaunchKernelGGL(hsic_pair_kernel, dim3(grid), dim3(kHsicThreads), lds, stream, pa, pb, seeds,
                     count, n, k, vecs, rank, offdiag, moment, permute, replicates,
                     include_observed, rb, out);
  return hipGetLastError();
}

hipError_t launch_dcc_rowmeans(const double* src_base, const long long* src_off,
                               const int* dst_slot, int count, int n, int stride, double index,
                               double* vecs, double* grand, int grid, hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  const size_t lds = (size_t)n * sizeof(double);
  if (index == 1.0)
    hipLaunchKernelGGL(dcc_rowmean_kernel<true>, dim3(grid), dim3(kHsicThreads), lds, stream,
                       src_base, src_off, dst_slot, count, n, stride, index, vecs, grand);
  else
    hipLaunchKernelGGL(dcc_rowmean_kernel<false>, dim3(grid), dim3(kHsicThreads), lds, stream,
                       src_base, src_off, dst_slot, count, n, stride, index, vecs, grand);
  return hipGetLastError();
}

hipError_t launch_dcc_perm_pairs(const int* pa, const int* pb, const unsigned long long* seeds,
                                 int count, int n, int stride, double index, const double* vecs,
                                 const double* grand, int replicates, int include_observed,
                                 double* out, int grid, hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  const int rb = perm_batch(n, replicates);
  const size_t lds = (size_t)3 * n * sizeof(double2) + (size_t)rb * n * sizeof(unsigned short);
  const void* fn = index == 1.0 ? (const void*)dcc_perm_kernel<true>
                                : (const void*)dcc_perm_kernel<false>;
  hipError_t err =
      hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (err != hipSuccess) return err;
  if (index == 1.0)
    hipLaunchKernelGGL(dcc_perm_kernel<true>, dim3(grid), dim3(kHsicThreads), lds, stream, pa, pb,
                       seeds, count, n, stride, index, vecs, grand, replicates, include_observed,
                       rb, out);
  else
    hipLaunchKernelGGL(dcc_perm_kernel<false>, dim3(grid), dim3(kHsicThreads), lds, stream, pa,
                       pb, seeds, count, n, stride, index, vecs, grand, replicates,
                       include_observed, rb, out);
  return hipGetLastError();
}

hipError_t launch_permutation_table(const unsigned long long* seeds, int count, int n,
                                    int replicates, int* out, hipStream_t stream) {
  const long long total = (long long)count * replicates;
  if (total <= 0) return hipSuccess;
  const int threads = 256;
  const long long blocks = (total + threads - 1) / threads;
  hipLaunchKernelGGL(permutation_table_kernel, dim3((unsigned)blocks), dim3(threads), 0, stream,
                     seeds, count, n, replicates, out);
  return hipGetLastError();
}

}  // namespace kpc
