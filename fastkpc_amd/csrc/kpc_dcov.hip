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
// One workgroup per pair: nV2 = sum_ij lx_i ly_j (Cx'Cy)_ij^2 / n, kpcalg
// moments and the gamma tail (fastkpc/src/legacy_dcov_gamma_cpp.cpp:420-455).
// dcov.gamma statistic of one test from the two components: with centred
// vectors X (n x k), Y (n x k) and values lx, ly,
// n V^2 = sum_ij lx_i ly_j (X_i . Y_j)^2 / n.  C = X'Y is a small GEMM: rows
// are staged through LDS in chunks of kPairRows (row-major, k padded to kp),
// each thread accumulates one or two 4 x 4 blocks of C in registers, and every
// C_ij is summed over the rows in order (as a plain dot product would).
constexpr int kPairThreads = 128;
constexpr int kPairRows = 32;  // 18 KB of LDS per workgroup at k = 35: eight per CU

__device__ __forceinline__ void pair_block_accumulate(const double* Xc, const double* Yc, int kp,
                                                      int rows, int bi, int bj, double* acc) {
  for (int r = 0; r < rows; ++r) {
    const double4 a = *reinterpret_cast<const double4*>(Xc + r * kp + 4 * bi);
    const double4 b = *reinterpret_cast<const double4*>(Yc + r * kp + 4 * bj);
    const double av[4] = {a.x, a.y, a.z, a.w};
    const double bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u * 4 + v] += av[u] * bv[v];
  }
}

__global__ __launch_bounds__(kPairThreads) void dcov_pair_kernel(
    const int* __restrict__ pair_x, const int* __restrict__ pair_y, int count, int n, int k,
    const double* __restrict__ comp_vals, const double* __restrict__ comp_vecs,
    const double* __restrict__ comp_dsum, const double* __restrict__ comp_moment,
    double* __restrict__ out) {
  extern __shared__ double lds_pair[];
  __shared__ double red[16];
  const int tid = threadIdx.x;
  const int kp = (k + 3) & ~3;
  const int nb = kp / 4;
  double* Xc = lds_pair;
  double* Yc = lds_pair + kPairRows * kp;
  const int b0 = tid, b1 = tid + (int)blockDim.x;
  const bool has0 = b0 < nb * nb, has1 = b1 < nb * nb;
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const int cx = pair_x[item], cy = pair_y[item];
    const double* X = comp_vecs + (size_t)cx * n * k;
    const double* Y = comp_vecs + (size_t)cy * n * k;
    const double* lx = comp_vals + (size_t)cx * k;
    const double* ly = comp_vals + (size_t)cy * k;
    double acc0[16], acc1[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      acc0[u] = 0.0;
      acc1[u] = 0.0;
    }
    for (int r0 = 0; r0 < n; r0 += kPairRows) {
      const int rows = min(kPairRows, n - r0);
      __syncthreads();
      // four elements of X and Y per thread in flight before any LDS store
      for (int e0 = tid; e0 < rows * kp; e0 += 4 * (int)blockDim.x) {
        double xv[4], yv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int e = e0 + u * (int)blockDim.x;
          const int i = e / rows, r = e % rows;
          const bool ok = e < rows * kp && i < k;
          xv[u] = ok ? X[(size_t)i * n + r0 + r] : 0.0;
          yv[u] = ok ? Y[(size_t)i * n + r0 + r] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int e = e0 + u * (int)blockDim.x;
          const int i = e / rows, r = e % rows;
          if (e < rows * kp) {
            Xc[r * kp + i] = xv[u];
            Yc[r * kp + i] = yv[u];
          }
        }
      }
      __syncthreads();
      if (has0) pair_block_accumulate(Xc, Yc, kp, rows, b0 % nb, b0 / nb, acc0);
      if (has1) pair_block_accumulate(Xc, Yc, kp, rows, b1 % nb, b1 / nb, acc1);
    }
    double part = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        if (has0) {
          const int i = 4 * (b0 % nb) + u, j = 4 * (b0 / nb) + v;
          if (i < k && j < k) part += lx[i] * ly[j] * acc0[u * 4 + v] * acc0[u * 4 + v];
        }
        if (has1) {
          const int i = 4 * (b1 % nb) + u, j = 4 * (b1 / nb) + v;
          if (i < k && j < k) part += lx[i] * ly[j] * acc1[u * 4 + v] * acc1[u * 4 + v];
        }
      }
    const double total = block_sum(part, red);
    if (tid == 0) {
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
  const int kp = (k + 3) & ~3;
  if ((kp / 4) * (kp / 4) > 2 * kPairThreads) return hipErrorInvalidValue;
  const size_t lds = (size_t)2 * kPairRows * kp * sizeof(double);
  hipLaunchKernelGGL(dcov_pair_kernel, dim3(grid), dim3(kPairThreads), lds, stream, px, py, count,
                     n, k, vals, vecs, dsum, moment, out);
  return hipGetLastError();
}

}  // namespace kpc
