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
// Pairs (hsic.gamma; one workgroup per test, strided):
//   HSIC = |Gcx^T Gcy|_F^2 / n^2; gamma moments and tail
//   (kpcalg/R/hsicgamma.R:74-83).  The permutation tests (hsic.perm, dcc.perm)
//   are in kpc_perm.hip; for hsic.perm the inchol kernel also writes the
//   centred factor row-major, zero-padded to 16 KT columns.
#include "kpc_common.h"
#include "kpc_device.h"
#include "kpc_gamma.h"

namespace kpc {

constexpr int kHsicThreads = 256;
constexpr int kRowsPerThread = kMaxN / kHsicThreads;
constexpr double kIncholTol = 0.001;

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
    int* __restrict__ rank_out, double* __restrict__ margin_out, double* __restrict__ rows,
    int kp) {
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
    // row-major copy of the centred factor, zero-padded to kp columns (the
    // operand layout of the hsic.perm replicate kernel)
    if (rows) {
      double* R = rows + (size_t)slot * n * kp;
      for (int e = tid; e < n * kp; e += blockDim.x) {
        const int i = e / kp, c = e - (e / kp) * kp;
        R[e] = c < k ? G[i + (size_t)n * c] : 0.0;
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------- hsic pairs
// Sum over this thread's (a, b) entries of (sum_i A[i,a] B[i,b])^2.
__device__ __forceinline__ double hsic_cross_part(const double* __restrict__ A, int ra,
                                                  const double* __restrict__ B, int rb, int n) {
  double part = 0.0;
  const int ne = ra * rb;
  for (int e = threadIdx.x; e < ne; e += blockDim.x) {
    const int a = e / rb, b = e - (e / rb) * rb;
    const double* ga = A + (size_t)n * a;
    const double* gb = B + (size_t)n * b;
    double dot = 0.0;
    for (int i = 0; i < n; ++i) dot += ga[i] * gb[i];
    part += dot * dot;
  }
  return part;
}

__global__ __launch_bounds__(kHsicThreads) void hsic_pair_kernel(
    const int* __restrict__ pa, const int* __restrict__ pb, int count, int n, int k,
    const double* __restrict__ vecs, const int* __restrict__ rank,
    const double* __restrict__ offdiag, const double* __restrict__ moment,
    double* __restrict__ out) {
  __shared__ double red[16];
  const int tid = threadIdx.x;
  const double nd = (double)n;
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const int ca = pa[item], cb = pb[item];
    const int ra = rank[ca], rb = rank[cb];
    const double* A = vecs + (size_t)ca * n * k;
    const double* B = vecs + (size_t)cb * n * k;
    const double observed = block_sum(hsic_cross_part(A, ra, B, rb, n), red) / (nd * nd);
    double* o = out + (size_t)item * kPairOut;
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

// ---------------------------------------------------------------- launchers
hipError_t launch_inchol_components(const double* src_base, const long long* src_off,
                                    const int* dst_slot, int count, int n, int k, double sigma,
                                    double* vecs, double* offdiag, double* moment, int* rank,
                                    double* margin, double* rows, int kp, int grid,
                                    hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  if (n > kMaxN || k > kMaxNumCol || (rows && kp < k)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(inchol_kernel, dim3(grid), dim3(kHsicThreads), 0, stream, src_base, src_off,
                     dst_slot, count, n, k, sigma, vecs, offdiag, moment, rank, margin, rows, kp);
  return hipGetLastError();
}

hipError_t launch_hsic_pairs(const int* pa, const int* pb, int count, int n, int k,
                             const double* vecs, const int* rank, const double* offdiag,
                             const double* moment, double* out, int grid, hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(hsic_pair_kernel, dim3(grid), dim3(kHsicThreads), 0, stream, pa, pb, count,
                     n, k, vecs, rank, offdiag, moment, out);
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

}  // namespace kpc
