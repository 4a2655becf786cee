// dcov.gamma components and pair statistics for gfx950.
//
// A component is the kpcalg dcov.gamma low-rank factor of one residual
// vector r (kpcalg/R/dcovgamma.R:67-75): the numCol eigenpairs of the distance
// matrix D_ij = |r_i - r_j| with the largest |lambda| (ties broken toward the
// higher ascending index, fastkpc/src/legacy_dcov_gamma_cpp.cpp:60-78), with
// the eigenvectors centred after selection (:135), plus sum(D) and the self
// moment sum_ij lambda_i lambda_j (C'C)_ij^2 (:283-300, :428-437).
//
// One workgroup owns one residual at a time and walks a strided list, so the
// grid size (and therefore the n x n scratch) is independent of the batch.
// The eigenpairs come from a Householder tridiagonalisation held in the
// workgroup's scratch slab (L2/MALL resident for n <= ~600), Sturm-count
// bisection for the 2k extreme eigenvalues, inverse iteration for the k
// selected eigenvectors (one lane each) and a back-transformation.
#include "kpc_common.h"
#include "kpc_device.h"
#include "kpc_gamma.h"

namespace kpc {

constexpr int kDcovThreads = 256;

// Number of eigenvalues of the tridiagonal (d, e) strictly below x.
__device__ int sturm_count(const double* d, const double* e2, int n, double x, double pivmin) {
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

__device__ double bisect_eigenvalue(const double* d, const double* e2, int n, int idx, double lo,
                                    double hi, double pivmin) {
  for (int it = 0; it < 200; ++it) {
    double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    double tol = 2.0 * 2.220446049250313e-16 * fmax(fabs(lo), fabs(hi));
    if (hi - lo <= tol + pivmin) break;
    if (sturm_count(d, e2, n, mid, pivmin) > idx) hi = mid; else lo = mid;
  }
  return 0.5 * (lo + hi);
}

// Inverse iteration for eigenvalue lam of (d, e); ws holds 4n doubles, z n.
__device__ void tridiag_invit(const double* d, const double* e, int n, double lam, double tnorm,
                              double* ws, double* z, unsigned seed) {
  double* dl = ws;
  double* dm = ws + n;
  double* du = ws + 2 * n;
  double* du2 = ws + 3 * n;
  const double pert = 2.220446049250313e-16 * tnorm;
  for (int i = 0; i < n; ++i) dm[i] = d[i] - lam;
  for (int i = 0; i < n - 1; ++i) { dl[i] = e[i]; du[i] = e[i]; }
  for (int i = 0; i < n; ++i) du2[i] = 0.0;
  // row-interchange flags live in the fifth n-slot of ws
  int* ip = reinterpret_cast<int*>(ws + 4 * n);
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
  unsigned s = seed * 2654435761u + 12345u;
  for (int i = 0; i < n; ++i) {
    s = s * 1103515245u + 12345u;
    z[i] = ((double)((s >> 8) & 0xFFFF) / 65536.0) - 0.5;
  }
  for (int it = 0; it < 4; ++it) {
    for (int i = 0; i < n - 1; ++i) {
      if (ip[i] == 0) {
        z[i + 1] -= dl[i] * z[i];
      } else {
        double t = z[i];
        z[i] = z[i + 1];
        z[i + 1] = t - dl[i] * z[i];
      }
    }
    z[n - 1] /= dm[n - 1];
    if (n >= 2) z[n - 2] = (z[n - 2] - du[n - 2] * z[n - 1]) / dm[n - 2];
    for (int i = n - 3; i >= 0; --i) z[i] = (z[i] - du[i] * z[i + 1] - du2[i] * z[i + 2]) / dm[i];
    double nrm = 0.0;
    for (int i = 0; i < n; ++i) nrm += z[i] * z[i];
    nrm = 1.0 / sqrt(nrm);
    for (int i = 0; i < n; ++i) z[i] *= nrm;
  }
}

// ------------------------------------------------------------ component
__global__ __launch_bounds__(kDcovThreads) void dcov_component_kernel(
    const double* __restrict__ src_base, const long long* __restrict__ src_off,
    const int* __restrict__ dst_slot, int count, int n, int k, double* __restrict__ comp_vals,
    double* __restrict__ comp_vecs, double* __restrict__ comp_dsum,
    double* __restrict__ comp_moment, double* __restrict__ scratch, long long scratch_stride) {
  extern __shared__ double lds_dyn[];
  const int nl = (n + 63) & ~63;
  double* r = lds_dyn;
  double* v = r + nl;
  double* w = v + nl;
  double* dg = w + nl;
  double* ofd = dg + nl;
  double* e2 = ofd + nl;
  double* red = e2 + nl;                 // 16
  double* cand_val = red + 16;           // 2 * kMaxNumCol
  double* sel_val = cand_val + 2 * kMaxNumCol;  // kMaxNumCol
  int* cand_idx = reinterpret_cast<int*>(sel_val + kMaxNumCol);  // 2 * kMaxNumCol
  double* A = scratch + (size_t)blockIdx.x * scratch_stride;
  double* tau = A + (size_t)n * n;           // n
  double* ivws = tau + n;                     // k * 5n
  const int tid = threadIdx.x;
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const double* x = src_base + src_off[item];
    const int slot = dst_slot[item];
    for (int i = tid; i < n; i += blockDim.x) r[i] = x[i];
    __syncthreads();
    // distance matrix, full storage, and its total in double-double
    dd acc{0.0, 0.0};
    for (int j = 0; j < n; ++j) {
      const double rj = r[j];
      for (int i = tid; i < n; i += blockDim.x) {
        double dv = fabs(r[i] - rj);
        A[(size_t)j * n + i] = dv;
        acc = dd_add(acc, dd{dv, 0.0});
      }
    }
    acc = block_sum_dd(acc, red);
    const double dsum = acc.hi + acc.lo;
    __syncthreads();
    // Householder tridiagonalisation
    for (int kk = 0; kk < n - 2; ++kk) {
      double part = 0.0;
      for (int i = kk + 2 + tid; i < n; i += blockDim.x) {
        double a = A[(size_t)kk * n + i];
        part += a * a;
      }
      const double xn2 = block_sum(part, red);
      const double alpha = A[(size_t)kk * n + kk + 1];
      double beta, tk, scal;
      if (xn2 == 0.0) {
        tk = 0.0;
        beta = alpha;
        scal = 0.0;
      } else {
        beta = -copysign(hypot(alpha, sqrt(xn2)), alpha);
        tk = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
      }
      if (tid == 0) {
        dg[kk] = A[(size_t)kk * n + kk];
        ofd[kk] = beta;
        tau[kk] = tk;
      }
      for (int i = kk + 1 + tid; i < n; i += blockDim.x) {
        double vi = (i == kk + 1) ? 1.0 : A[(size_t)kk * n + i] * scal;
        v[i] = vi;
        if (i >= kk + 2) A[(size_t)kk * n + i] = vi;
      }
      __syncthreads();
      if (tk != 0.0) {
        // p = tau * A22 v
        double pv_part = 0.0;
        for (int i = kk + 1 + tid; i < n; i += blockDim.x) {
          double s = 0.0;
          for (int j = kk + 1; j < n; ++j) s += A[(size_t)j * n + i] * v[j];
          s *= tk;
          w[i] = s;
          pv_part += s * v[i];
        }
        const double pv = block_sum(pv_part, red);
        const double half = -0.5 * tk * pv;
        for (int i = kk + 1 + tid; i < n; i += blockDim.x) w[i] += half * v[i];
        __syncthreads();
        for (int j = kk + 1; j < n; ++j) {
          const double vj = v[j], wj = w[j];
          for (int i = kk + 1 + tid; i < n; i += blockDim.x)
            A[(size_t)j * n + i] -= v[i] * wj + w[i] * vj;
        }
      }
      __syncthreads();
    }
    if (tid == 0) {
      if (n >= 2) {
        dg[n - 2] = A[(size_t)(n - 2) * n + n - 2];
        ofd[n - 2] = A[(size_t)(n - 2) * n + n - 1];
        tau[n - 2] = 0.0;
      }
      dg[n - 1] = A[(size_t)(n - 1) * n + n - 1];
    }
    __syncthreads();
    for (int i = tid; i < n - 1; i += blockDim.x) e2[i] = ofd[i] * ofd[i];
    // Gershgorin bounds and pivmin
    double gl_part = 1e308, gu_part = -1e308, tn_part = 0.0, e2max_part = 0.0;
    for (int i = tid; i < n; i += blockDim.x) {
      double rad = (i > 0 ? fabs(ofd[i - 1]) : 0.0) + (i < n - 1 ? fabs(ofd[i]) : 0.0);
      gl_part = fmin(gl_part, dg[i] - rad);
      gu_part = fmax(gu_part, dg[i] + rad);
      tn_part = fmax(tn_part, fabs(dg[i]) + rad);
      if (i < n - 1) e2max_part = fmax(e2max_part, ofd[i] * ofd[i]);
    }
    const double gl = -block_max(-gl_part, red);
    const double gu = block_max(gu_part, red);
    const double tnorm = block_max(tn_part, red);
    const double e2max = block_max(e2max_part, red);
    const double pivmin = 2.2250738585072014e-308 * fmax(1.0, e2max);
    const double span = gu - gl;
    const double lo0 = gl - 2.0 * 2.220446049250313e-16 * span - pivmin;
    const double hi0 = gu + 2.0 * 2.220446049250313e-16 * span + pivmin;
    // candidate ascending indices: {0..k-1} U {max(k, n-k)..n-1}
    const int nlow = k < n ? k : n;
    const int hi_start = (n - k) > nlow ? (n - k) : nlow;
    const int ncand = nlow + (n - hi_start);
    for (int c = tid; c < ncand; c += blockDim.x) {
      const int idx = c < nlow ? c : hi_start + (c - nlow);
      cand_idx[c] = idx;
      cand_val[c] = bisect_eigenvalue(dg, e2, n, idx, lo0, hi0, pivmin);
    }
    __syncthreads();
    if (tid == 0) {
      // selection: |lambda| descending, tie -> higher ascending index
      for (int s = 0; s < k; ++s) {
        int best = -1;
        for (int c = 0; c < ncand; ++c) {
          if (cand_idx[c] < 0) continue;
          if (best < 0) { best = c; continue; }
          double a = fabs(cand_val[c]), b = fabs(cand_val[best]);
          if (a > b || (a == b && cand_idx[c] > cand_idx[best])) best = c;
        }
        sel_val[s] = cand_val[best];
        cand_idx[best] = -1;
      }
    }
    __syncthreads();
    double* Z = comp_vecs + (size_t)slot * n * k;
    for (int c = tid; c < k; c += blockDim.x)
      tridiag_invit(dg, ofd, n, sel_val[c], tnorm, ivws + (size_t)c * 5 * n, Z + (size_t)c * n,
                    (unsigned)c);
    __syncthreads();
    // back-transformation Z <- H_0 ... H_{n-3} Z
    const int lane = tid & (KPC_WAVE - 1);
    const int wid = tid / KPC_WAVE;
    const int nw = blockDim.x / KPC_WAVE;
    for (int rr = n - 3; rr >= 0; --rr) {
      const double t = tau[rr];
      if (t == 0.0) continue;
      for (int i = rr + 1 + tid; i < n; i += blockDim.x) v[i] = (i == rr + 1) ? 1.0 : A[(size_t)rr * n + i];
      __syncthreads();
      for (int c = wid; c < k; c += nw) {
        double* z = Z + (size_t)c * n;
        double s = 0.0;
        for (int i = rr + 1 + lane; i < n; i += KPC_WAVE) s += v[i] * z[i];
        s = wave_sum(s) * t;
        for (int i = rr + 1 + lane; i < n; i += KPC_WAVE) z[i] -= s * v[i];
      }
      __syncthreads();
    }
    // centre columns (plain mean, as arma::mean)
    for (int c = wid; c < k; c += nw) {
      double* z = Z + (size_t)c * n;
      double s = 0.0;
      for (int i = lane; i < n; i += KPC_WAVE) s += z[i];
      s = wave_sum(s) / (double)n;
      for (int i = lane; i < n; i += KPC_WAVE) z[i] -= s;
    }
    __syncthreads();
    // self moment sum_ij l_i l_j (C'C)_ij^2
    double mpart = 0.0;
    for (int e = tid; e < k * k; e += blockDim.x) {
      const int i = e % k, j = e / k;
      const double* ci = Z + (size_t)i * n;
      const double* cj = Z + (size_t)j * n;
      double s = 0.0;
      for (int rr = 0; rr < n; ++rr) s += ci[rr] * cj[rr];
      mpart += sel_val[i] * sel_val[j] * s * s;
    }
    const double moment = block_sum(mpart, red);
    if (tid == 0) {
      comp_dsum[slot] = dsum;
      comp_moment[slot] = moment;
    }
    for (int c = tid; c < k; c += blockDim.x) comp_vals[(size_t)slot * k + c] = sel_val[c];
    __syncthreads();
  }
}

// ------------------------------------------------------------------ pair
// dcov.gamma statistic of one test from the two components (kpcalg moments
// and the gamma tail, fastkpc/src/legacy_dcov_gamma_cpp.cpp:420-455): with
// centred vectors X, Y (n x k, column-major) and values lx, ly,
// n V^2 = sum_ij lx_i ly_j (X_i . Y_j)^2 / n.  C = X'Y is a k x k GEMM over the
// n samples, run on FP64 matrix cores: one wave per test, C in KT x KT tiles
// of 16 x 16 (k padded to 16 KT with zeros), v_mfma_f64_16x16x4f64 over four
// samples per step.  Operand fragments: lane l holds X[16 a + l % 16][r + l / 16]
// (A, 16 components x 4 samples) and Y[16 b + l % 16][r + l / 16] (B); item v of
// the accumulator is C[16 a + 4 v + l / 16][16 b + l % 16].  Four samples' loads
// run a step ahead of the matrix ops; no LDS, no barrier.
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int kPairWaves = 4;  // tests per 256-lane workgroup

template <int KT>
__device__ __forceinline__ void pair_fragments(const double* X, const double* Y, int n, int k,
                                               int r, int c, double (&av)[KT], double (&bv)[KT]) {
  const bool rv = r < n;
  const int rr = rv ? r : 0;
#pragma unroll
  for (int t = 0; t < KT; ++t) {
    const int i = 16 * t + c;
    const bool ok = rv && i < k;
    const int ii = i < k ? i : 0;
    const double x = X[(size_t)ii * n + rr];
    const double y = Y[(size_t)ii * n + rr];
    av[t] = ok ? x : 0.0;
    bv[t] = ok ? y : 0.0;
  }
}

template <int KT>
__global__ __launch_bounds__(64 * kPairWaves) void dcov_pair_kernel(
    const int* __restrict__ pair_x, const int* __restrict__ pair_y, int count, int n, int k,
    const double* __restrict__ comp_vals, const double* __restrict__ comp_vecs,
    const double* __restrict__ comp_dsum, const double* __restrict__ comp_moment,
    double* __restrict__ out) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int wid = threadIdx.x / KPC_WAVE;
  const int c = lane & 15, g = lane >> 4;
  for (int item = blockIdx.x * kPairWaves + wid; item < count; item += gridDim.x * kPairWaves) {
    const int cx = pair_x[item], cy = pair_y[item];
    const double* X = comp_vecs + (size_t)cx * n * k;
    const double* Y = comp_vecs + (size_t)cy * n * k;
    const double* lx = comp_vals + (size_t)cx * k;
    const double* ly = comp_vals + (size_t)cy * k;
    v4d acc[KT][KT];
#pragma unroll
    for (int a = 0; a < KT; ++a)
#pragma unroll
      for (int b = 0; b < KT; ++b) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
    // two 4-row steps per trip with the next two steps' loads in flight; the
    // accumulation runs over the rows in order (a step past n adds exact
    // zeros: its operands are zero-padded)
    double av[KT], bv[KT], aw[KT], bw[KT], an[KT], bn[KT], ao[KT], bo[KT];
    pair_fragments<KT>(X, Y, n, k, g, c, av, bv);
    pair_fragments<KT>(X, Y, n, k, 4 + g, c, aw, bw);
    for (int r0 = 0; r0 < n; r0 += 8) {
      pair_fragments<KT>(X, Y, n, k, r0 + 8 + g, c, an, bn);
      pair_fragments<KT>(X, Y, n, k, r0 + 12 + g, c, ao, bo);
#pragma unroll
      for (int a = 0; a < KT; ++a)
#pragma unroll
        for (int b = 0; b < KT; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
#pragma unroll
      for (int a = 0; a < KT; ++a)
#pragma unroll
        for (int b = 0; b < KT; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(aw[a], bw[b], acc[a][b], 0, 0, 0);
#pragma unroll
      for (int t = 0; t < KT; ++t) {
        av[t] = an[t];
        bv[t] = bn[t];
        aw[t] = ao[t];
        bw[t] = bo[t];
      }
    }
    double part = 0.0;
#pragma unroll
    for (int a = 0; a < KT; ++a)
#pragma unroll
      for (int b = 0; b < KT; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = 16 * a + 4 * v + g, j = 16 * b + c;
          if (i < k && j < k) part += lx[i] * ly[j] * acc[a][b][v] * acc[a][b][v];
        }
    const double total = wave_sum(part);
    if (lane == 0) {
      const double nd = (double)n;
      const double nV2 = total / nd;
      const double mean = (comp_dsum[cx] / (nd * nd)) * (comp_dsum[cy] / (nd * nd));
      const double vf = 2.0 * (nd - 4.0) * (nd - 5.0) / nd / (nd - 1.0) / (nd - 2.0) / (nd - 3.0);
      const double var = vf * comp_moment[cx] * comp_moment[cy] / pow(nd, 4.0) * pow(nd, 2.0);
      double* o = out + (size_t)item * kPairOut;
      o[0] = gamma_pvalue(nV2, mean, var);
      o[1] = nV2;
      o[2] = mean;
      o[3] = var;
      o[4] = 1e300;
    }
  }
}

// ------------------------------------------------------------ launchers
size_t dcov_component_lds_bytes(int n) {
  const int nl = (n + 63) & ~63;
  return (size_t)(6 * nl + 16 + 3 * kMaxNumCol) * sizeof(double) + 2 * kMaxNumCol * sizeof(int);
}

long long dcov_component_scratch_doubles(int n, int k) {
  long long s = (long long)n * n + n + (long long)k * 5 * n + 8;
  return (s + 31) / 32 * 32;
}

hipError_t launch_dcov_components(const double* src_base, const long long* src_off,
                                  const int* dst_slot, int count, int n, int k, double* vals,
                                  double* vecs, double* dsum, double* moment, double* scratch,
                                  int grid, hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  const long long stride = dcov_component_scratch_doubles(n, k);
  const size_t lds = dcov_component_lds_bytes(n);
  hipError_t err = hipFuncSetAttribute((const void*)dcov_component_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(dcov_component_kernel, dim3(grid), dim3(kDcovThreads), lds, stream, src_base,
                     src_off, dst_slot, count, n, k, vals, vecs, dsum, moment, scratch, stride);
  return hipGetLastError();
}

hipError_t launch_dcov_pairs(const int* px, const int* py, int count, int n, int k,
                             const double* vals, const double* vecs, const double* dsum,
                             const double* moment, double* out, int grid, hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  if (k < 1 || k > 64) return hipErrorInvalidValue;
  // grid: at most one workgroup per kPairWaves tests
  const int g = std::max(1, std::min(grid, (count + kPairWaves - 1) / kPairWaves));
  const dim3 block(64 * kPairWaves);
  switch ((k + 15) / 16) {
    case 1:
      hipLaunchKernelGGL(dcov_pair_kernel<1>, dim3(g), block, 0, stream, px, py, count, n, k, vals,
                         vecs, dsum, moment, out);
      break;
    case 2:
      hipLaunchKernelGGL(dcov_pair_kernel<2>, dim3(g), block, 0, stream, px, py, count, n, k, vals,
                         vecs, dsum, moment, out);
      break;
    case 3:
      hipLaunchKernelGGL(dcov_pair_kernel<3>, dim3(g), block, 0, stream, px, py, count, n, k, vals,
                         vecs, dsum, moment, out);
      break;
    default:
      hipLaunchKernelGGL(dcov_pair_kernel<4>, dim3(g), block, 0, stream, px, py, count, n, k, vals,
                         vecs, dsum, moment, out);
      break;
  }
  return hipGetLastError();
}

}  // namespace kpc
