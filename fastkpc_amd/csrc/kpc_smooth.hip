// Thin-plate regression smooth construction (mgcv 1.9-1 `tp`, k = 10 for one
// covariate, 30 for two) on gfx950: one workgroup builds one smooth.
//
// Reference semantics: fastkpc/src/full_cuda_ci_native_setup.cpp
//   :96-107  shift = R mean of each covariate (long double + correction)
//   :114-152 eta constants, :154-198 null space / polynomial powers
//   :200-239 lexicographically sorted unique locations
//   :303-494 Rlanczos: LCG start vector, Lanczos with two reorthogonalisation
//            passes, Ritz check every max(10, k/2) (capped at n/10) steps
//            once the step index reaches k, eps^0.7 * ||T|| stop rule with the
//            lower-end index quirk at :439, Ritz vector assembly
//   :496-594 QT Householders, hq_post, hq_pre_transpose
//   :621-688 sum-to-zero constraint (LINPACK dqrdc2/dqrqty, one reflector)
//   :690-858 X = [U D Z | T], penalty Z' D Z, RMS column scaling,
//            S <- S / (||S||_1 / ||X||_inf^2)
// Reorthogonalisation is done as two classical Gram-Schmidt passes (the
// reference uses modified Gram-Schmidt); both give orthogonality to working
// precision.  Ritz values/vectors of the Lanczos tridiagonal come from
// implicit QL whose Givens sequences are computed by one lane and applied to
// the eigenvector rows by the whole workgroup.
#include "kpc_common.h"
#include "kpc_device.h"

namespace kpc {

constexpr int kSmoothThreads = 256;
constexpr int kMaxBasis = 32;

__device__ double eta_constant(int m, int d) {
  const double pi = asin(1.0) * 2.0;
  const double gamma_half = sqrt(pi);
  const int d2 = d / 2;
  const int m2 = 2 * m;
  double value = 1.0;
  if (d % 2 == 0) {
    value = ((m + 1 + d2) % 2) ? -1.0 : 1.0;
    for (int i = 0; i < m2 - 1; ++i) value /= 2.0;
    for (int i = 0; i < d2; ++i) value /= pi;
    for (int i = 2; i < m; ++i) value /= (double)i;
    for (int i = 2; i <= m - d2; ++i) value /= (double)i;
  } else {
    value = gamma_half;
    const int kk = m - (d - 1) / 2;
    for (int i = 0; i < kk; ++i) value /= -0.5 - (double)i;
    for (int i = 0; i < m; ++i) value /= 4.0;
    for (int i = 0; i < d2; ++i) value /= pi;
    value /= gamma_half;
    for (int i = 2; i < m; ++i) value /= (double)i;
  }
  return value;
}

__device__ double eta_r2(int m, int d, double r2, double eta0) {
  if (r2 <= 0.0) return 0.0;
  const int d2 = d / 2;
  double value = eta0;
  if (d % 2 == 0) {
    value *= log(r2) * 0.5;
    for (int i = 0; i < m - d2; ++i) value *= r2;
  } else {
    for (int i = 0; i < m - d2 - 1; ++i) value *= r2;
    value *= sqrt(r2);
  }
  return value;
}

// Workgroup LDS, carved from dynamic shared memory with arrays of np2 (the
// next power of two >= n) entries.
struct SmoothLds {
  double* ka;   // sort key 0 -> unique coordinate 0
  double* kb;   // sort key 1 -> unique coordinate 1
  int* ki;      // sort index -> original row -> unique id map
  double* z;    // Lanczos residual; Givens cos; QT reflectors
  double* al;   // Lanczos alpha; QL diagonal
  double* be;   // Lanczos beta
  double* cc;   // CGS coefficients; QL off-diagonal; column means
  double* ss;   // Givens sin
  double* tv;   // Ritz values, descending
  int* perm;    // descending order -> QL column
  double* red;  // 16
  int* ctl;     // 8
  double* dctl; // 8
  double* zq;   // 3 * kMaxBasis
  double* cm;   // kMaxBasis
  double* rms;  // kMaxBasis
};

__host__ __device__ inline size_t smooth_lds_size(int np2) {
  return (size_t)np2 * (8 * sizeof(double) + 2 * sizeof(int)) +
         (16 + 8 + 3 * kMaxBasis + 2 * kMaxBasis) * sizeof(double) + 8 * sizeof(int);
}

__device__ inline SmoothLds carve_smooth_lds(double* base, int np2) {
  SmoothLds L;
  double* d = base;
  L.ka = d; d += np2;
  L.kb = d; d += np2;
  L.z = d; d += np2;
  L.al = d; d += np2;
  L.be = d; d += np2;
  L.cc = d; d += np2;
  L.ss = d; d += np2;
  L.tv = d; d += np2;
  L.red = d; d += 16;
  L.dctl = d; d += 8;
  L.zq = d; d += 3 * kMaxBasis;
  L.cm = d; d += kMaxBasis;
  L.rms = d; d += kMaxBasis;
  int* ip = reinterpret_cast<int*>(d);
  L.ki = ip; ip += np2;
  L.perm = ip; ip += np2;
  L.ctl = ip;
  return L;
}

// Implicit QL (EISPACK tql2 ordering) on LDS (d, e) with eigenvectors in the
// global size x size matrix Zm (column-major).  Lane 0 generates each sweep's
// Givens rotations, every lane applies them to its rows.
__device__ void ql_eigen_parallel(const SmoothLds& L, int size, double* Zm) {
  const int tid = threadIdx.x;
  double* d = L.al;
  double* e = L.cc;  // e[i] = T(i, i+1), e[size-1] = 0
  for (int i = tid; i < size * size; i += blockDim.x) Zm[i] = 0.0;
  __syncthreads();
  for (int i = tid; i < size; i += blockDim.x) Zm[(size_t)i * size + i] = 1.0;
  if (tid == 0) { L.ctl[0] = 0; L.ctl[1] = 0; L.ctl[4] = 0; }
  __syncthreads();
  while (true) {
    if (tid == 0) {
      int l = L.ctl[0];
      int found = 0;
      while (l < size && !found) {
        int m;
        for (m = l; m < size - 1; ++m) {
          double dd2 = fabs(d[m]) + fabs(d[m + 1]);
          if (fabs(e[m]) <= 2.220446049250313e-16 * dd2) break;
        }
        if (m == l) { ++l; L.ctl[4] = 0; continue; }
        if (L.ctl[4]++ >= 60) { ++l; L.ctl[4] = 0; continue; }
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + copysign(r, g));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        int early = 0;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          e[i + 1] = (r = hypot(f, g));
          if (r == 0.0) { d[i + 1] -= p; e[m] = 0.0; early = 1; break; }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          d[i + 1] = g + (p = s * r);
          g = c * r - b;
          L.z[i] = c;     // rotation (c, s) acting on columns (i, i+1)
          L.ss[i] = s;
        }
        if (!early) {
          d[l] -= p;
          e[l] = g;
          e[m] = 0.0;
        }
        L.ctl[1] = m - 1;          // first rotation index
        L.ctl[2] = early ? i + 1 : l;  // last rotation index
        found = 1;
      }
      L.ctl[0] = l;
      L.ctl[3] = found;
    }
    __syncthreads();
    if (!L.ctl[3]) break;
    const int ihi = L.ctl[1], ilo = L.ctl[2];
    for (int k = tid; k < size; k += blockDim.x) {
      for (int i = ihi; i >= ilo; --i) {
        const double c = L.z[i], s = L.ss[i];
        double* zi = Zm + (size_t)i * size + k;
        double* zi1 = Zm + (size_t)(i + 1) * size + k;
        const double f = *zi1;
        *zi1 = s * *zi + c * f;
        *zi = c * *zi - s * f;
      }
    }
    __syncthreads();
  }
  __syncthreads();
}

__global__ __launch_bounds__(kSmoothThreads) void smooth_setup_kernel(
    const double* __restrict__ data, int n, int p, const int* __restrict__ items, int count,
    double* __restrict__ out_X, double* __restrict__ out_S, int* __restrict__ out_info,
    double* __restrict__ scratch, long long scratch_stride) {
  extern __shared__ double lds_raw[];
  int np2 = 1;
  while (np2 < n) np2 <<= 1;
  SmoothLds L = carve_smooth_lds(lds_raw, np2);
  const int tid = threadIdx.x;
  const int lane = tid & (KPC_WAVE - 1);
  const int wid = tid / KPC_WAVE;
  const int nw = blockDim.x / KPC_WAVE;
  double* ws = scratch + (size_t)blockIdx.x * scratch_stride;
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const int d = items[4 * item + 0];
    const int c0 = items[4 * item + 1];
    const int c1 = items[4 * item + 2];
    const int slot = items[4 * item + 3];
    const int kb = d == 1 ? 10 : 30;
    const int M = d == 1 ? 2 : 3;
    const int order = 2;
    int* info = out_info + 4 * slot;
    // ---- shifted coordinates as sort keys
    const double* col0 = data + (size_t)c0 * n;
    const double* col1 = data + (size_t)c1 * n;
    const double mu0 = block_r_mean(col0, n, L.red);
    double mu1 = 0.0;
    if (d == 2) mu1 = block_r_mean(col1, n, L.red);
    for (int i = tid; i < np2; i += blockDim.x) {
      if (i < n) {
        L.ka[i] = col0[i] - mu0;
        L.kb[i] = d == 2 ? col1[i] - mu1 : 0.0;
        L.ki[i] = i;
      } else {
        L.ka[i] = __builtin_inf();
        L.kb[i] = __builtin_inf();
        L.ki[i] = 0x7fffffff;
      }
    }
    __syncthreads();
    // ---- bitonic sort on (ka, kb, ki)
    for (int size = 2; size <= np2; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int t = tid; t < np2 / 2; t += blockDim.x) {
          const int lo = 2 * t - (t & (stride - 1));
          const int hi = lo + stride;
          const bool up = ((lo & size) == 0);
          const double a0 = L.ka[lo], b0 = L.ka[hi];
          const double a1 = L.kb[lo], b1 = L.kb[hi];
          const int ai = L.ki[lo], bi = L.ki[hi];
          bool greater = (a0 > b0) || (a0 == b0 && (a1 > b1 || (a1 == b1 && ai > bi)));
          if (greater == up) {
            L.ka[lo] = b0; L.ka[hi] = a0;
            L.kb[lo] = b1; L.kb[hi] = a1;
            L.ki[lo] = bi; L.ki[hi] = ai;
          }
        }
        __syncthreads();
      }
    }
    // ---- unique ids (serial scan by lane 0 over <= 2048 entries)
    if (tid == 0) {
      int nu = 0;
      for (int t = 0; t < n; ++t) {
        bool is_new = t == 0 || L.ka[t] != L.ka[t - 1] || (d == 2 && L.kb[t] != L.kb[t - 1]);
        if (is_new) {
          L.ka[nu] = L.ka[t];
          L.kb[nu] = L.kb[t];
          ++nu;
        }
        L.z[t] = (double)(nu - 1);  // unique id of sorted position t
      }
      L.ctl[5] = nu;
    }
    __syncthreads();
    const int nu = L.ctl[5];
    // map original row -> unique id (ki holds original rows in sorted order)
    for (int t = tid; t < n; t += blockDim.x) {
      const int orig = L.ki[t];
      ws[orig] = L.z[t];  // temporary: map in scratch head
    }
    __syncthreads();
    for (int i = tid; i < n; i += blockDim.x) L.ki[i] = (int)ws[i];
    __syncthreads();
    if (nu < kb || nu > 2000) {
      if (tid == 0) { info[0] = -3; info[1] = nu; }
      __syncthreads();
      continue;
    }
    // ---- scratch layout
    double* E = ws;                                  // nu x nu
    double* Qb = E + (size_t)nu * nu;                // nu x (nu+1)
    double* Zm = Qb + (size_t)nu * (nu + 1);         // up to nu x nu (QL vectors)
    double* U = Zm + (size_t)nu * nu;                // nu x kb
    double* Xs = U + (size_t)nu * kb;                // n x kb (row-major) before absorption
    const double eta0 = eta_constant(order, d);
    for (int j = 0; j < nu; ++j) {
      for (int i = tid; i < nu; i += blockDim.x) {
        double s2;
        double dx = L.ka[i] - L.ka[j];
        s2 = dx * dx;
        if (d == 2) {
          double dy = L.kb[i] - L.kb[j];
          s2 += dy * dy;
        }
        E[(size_t)j * nu + i] = (i == j) ? 0.0 : eta_r2(order, d, s2, eta0);
      }
    }
    // ---- Lanczos start vector (mgcv LCG)
    if (tid == 0) {
      unsigned long st = 1;
      double nrm = 0.0;
      for (int r = 0; r < nu; ++r) {
        st = (st * 106UL + 1283UL) % 6075UL;
        double vv = (double)st / 6075.0 - 0.5;
        Qb[r] = vv;
        nrm += vv * vv;
      }
      L.dctl[0] = sqrt(nrm);
    }
    __syncthreads();
    for (int r = tid; r < nu; r += blockDim.x) Qb[r] /= L.dctl[0];
    __syncthreads();
    const double tol = pow(2.220446049250313e-16, 0.7);
    int check = kb / 2;
    if (check < 10) check = 10;
    int tenth = (int)floor(nu / 10.0);
    if (tenth < 1) tenth = 1;
    if (tenth < check) check = tenth;
    int upper = kb, lower = 0, completed = nu, tri_len = 0, converged = 0;
    for (int it = 0; it < nu; ++it) {
      const double* qi = Qb + (size_t)it * nu;
      for (int r = tid; r < nu; r += blockDim.x) {
        double s = 0.0;
        for (int c = 0; c < nu; ++c) s += E[(size_t)c * nu + r] * qi[c];
        L.z[r] = s;
      }
      double part = 0.0;
      for (int r = tid; r < nu; r += blockDim.x) part += qi[r] * L.z[r];
      const double proj = block_sum(part, L.red);
      if (tid == 0) L.al[it] = proj;
      if (it == 0) {
        for (int r = tid; r < nu; r += blockDim.x) L.z[r] -= proj * Qb[r];
      } else {
        const double prior = L.be[it - 1];
        const double* qp = Qb + (size_t)(it - 1) * nu;
        for (int r = tid; r < nu; r += blockDim.x) L.z[r] -= proj * qi[r] + prior * qp[r];
        __syncthreads();
        for (int pass = 0; pass < 2; ++pass) {
          for (int b = wid; b <= it; b += nw) {
            const double* qb = Qb + (size_t)b * nu;
            double s = 0.0;
            for (int r = lane; r < nu; r += KPC_WAVE) s += L.z[r] * qb[r];
            s = wave_sum(s);
            if (lane == 0) L.cc[b] = s;
          }
          __syncthreads();
          for (int r = tid; r < nu; r += blockDim.x) {
            double s = 0.0;
            for (int b = 0; b <= it; ++b) s += L.cc[b] * Qb[(size_t)b * nu + r];
            L.z[r] -= s;
          }
          __syncthreads();
        }
      }
      __syncthreads();
      part = 0.0;
      for (int r = tid; r < nu; r += blockDim.x) part += L.z[r] * L.z[r];
      const double rn = sqrt(block_sum(part, L.red));
      if (tid == 0) L.be[it] = rn;
      if (it < nu - 1) {
        double* qn = Qb + (size_t)(it + 1) * nu;
        for (int r = tid; r < nu; r += blockDim.x) qn[r] = L.z[r] / rn;
      }
      __syncthreads();
      if ((it >= kb && it % check == 0) || it == nu - 1) {
        const int size = it + 1;
        // QL on a copy of (alpha, beta)
        // al is destroyed by QL, so keep the Lanczos alpha in tv temporarily
        for (int i = tid; i < size; i += blockDim.x) {
          L.tv[i] = L.al[i];
          L.cc[i] = (i < size - 1) ? L.be[i] : 0.0;
        }
        __syncthreads();
        ql_eigen_parallel(L, size, Zm);
        // order: descending eigenvalues; restore alpha
        if (tid == 0) {
          for (int i = 0; i < size; ++i) L.perm[i] = i;
          for (int i = 1; i < size; ++i) {
            int pi = L.perm[i];
            int j = i - 1;
            while (j >= 0 && L.al[L.perm[j]] < L.al[pi]) { L.perm[j + 1] = L.perm[j]; --j; }
            L.perm[j + 1] = pi;
          }
        }
        __syncthreads();
        for (int i = tid; i < size; i += blockDim.x) {
          const double lam = L.al[L.perm[i]];
          L.ss[i] = lam;  // descending Ritz values (temporary)
        }
        __syncthreads();
        for (int i = tid; i < size; i += blockDim.x) {
          const double a = L.tv[i];
          L.tv[i] = L.ss[i];
          L.al[i] = a;
        }
        __syncthreads();
        if (tid == 0) {
          double tnorm = fabs(L.tv[0]);
          if (fabs(L.tv[it]) > tnorm) tnorm = fabs(L.tv[it]);
          int conv = 0, pos = 0, neg = 0;
          if (it >= kb) {
            const double maxerr = tnorm * tol;
            conv = 1;
            while (pos + neg < kb) {
              if (fabs(L.tv[pos]) >= fabs(L.tv[it - neg])) {
                double err = fabs(rn * Zm[(size_t)L.perm[pos] * size + it]);
                if (err > maxerr) { conv = 0; break; }
                ++pos;
              } else {
                double err = fabs(rn * Zm[(size_t)L.perm[neg] * size + it]);  // index quirk
                if (err > maxerr) { conv = 0; break; }
                ++neg;
              }
            }
          }
          L.ctl[6] = conv;
          L.ctl[7] = pos;
          L.ctl[5] = neg;
        }
        __syncthreads();
        tri_len = size;
        if (L.ctl[6]) {
          upper = L.ctl[7];
          lower = L.ctl[5];
          completed = it + 1;
          converged = 1;
        }
        __syncthreads();
        if (converged) break;
      }
    }
    (void)converged;
    // ---- Ritz vectors U (nu x kb) and values
    for (int e = 0; e < upper + lower; ++e) {
      const int src = e < upper ? e : completed - (upper + lower - e);
      const double* sv = Zm + (size_t)L.perm[src] * tri_len;
      for (int r = tid; r < nu; r += blockDim.x) {
        double s = 0.0;
        for (int b = 0; b < completed; ++b) s += Qb[(size_t)b * nu + r] * sv[b];
        U[(size_t)e * nu + r] = s;
      }
      if (tid == 0) L.cm[e] = L.tv[src];
    }
    __syncthreads();
    // ---- T columns [1, x (, y)]; TU = T' U (M x kb)
    for (int e = wid; e < M * kb; e += nw) {
      const int row = e / kb, col = e % kb;
      double s = 0.0;
      for (int r = lane; r < nu; r += KPC_WAVE) {
        double t = row == 0 ? 1.0 : (row == 1 ? L.ka[r] : L.kb[r]);
        s += t * U[(size_t)col * nu + r];
      }
      s = wave_sum(s);
      if (lane == 0) L.zq[row * kb + col] = s;
    }
    __syncthreads();
    if (tid == 0) {
      // QT householders: zq (input, row-major) -> stored reflectors in z[]
      double* in = L.zq;
      double* st = L.z;  // M x kb
      const int cols = kb;
      for (int step = 0; step < M; ++step) {
        double* rs = in + step * cols;
        double scale = 0.0;
        for (int c = 0; c < cols - step; ++c) if (fabs(rs[c]) > scale) scale = fabs(rs[c]);
        if (scale != 0.0) for (int c = 0; c < cols - step; ++c) rs[c] /= scale;
        double len = 0.0;
        for (int c = 0; c < cols - step; ++c) len += rs[c] * rs[c];
        len = sqrt(len);
        if (rs[cols - step - 1] < 0.0) len = -len;
        rs[cols - step - 1] += len;
        double mult = len != 0.0 ? 1.0 / (len * rs[cols - step - 1]) : 0.0;
        double terminal = len * scale;
        for (int r = step + 1; r < M; ++r) {
          double* rr = in + r * cols;
          double pr = 0.0;
          for (int c = 0; c < cols - step; ++c) pr += rs[c] * rr[c];
          pr *= mult;
          for (int c = 0; c < cols - step; ++c) rr[c] -= pr * rs[c];
        }
        double rm = sqrt(mult);
        for (int c = 0; c < cols - step; ++c) st[step * cols + c] = rs[c] * rm;
        for (int c = cols - step; c < cols; ++c) st[step * cols + c] = 0.0;
        rs[cols - step - 1] = -terminal;
        for (int c = 0; c < cols - step - 1; ++c) rs[c] = 0.0;
      }
    }
    __syncthreads();
    // ---- X rows: unique row u = U D Z..., then T columns; expand to n rows
    for (int r = tid; r < n; r += blockDim.x) {
      const int u = L.ki[r];
      double row[kMaxBasis];
      for (int c = 0; c < kb; ++c) row[c] = U[(size_t)c * nu + u] * L.cm[c];
      for (int t = 0; t < M; ++t) {
        const double* ht = L.z + t * kb;
        double pr = 0.0;
        for (int c = 0; c < kb; ++c) pr += row[c] * ht[c];
        for (int c = 0; c < kb; ++c) row[c] -= pr * ht[c];
      }
      row[kb - M] = 1.0;
      row[kb - M + 1] = L.ka[u];
      if (M == 3) row[kb - 1] = L.kb[u];
      for (int c = 0; c < kb; ++c) Xs[(size_t)r * kb + c] = row[c];
    }
    __syncthreads();
    // ---- RMS column scaling
    for (int c = wid; c < kb; c += nw) {
      double s = 0.0;
      for (int r = lane; r < n; r += KPC_WAVE) {
        double x = Xs[(size_t)r * kb + c];
        s += x * x;
      }
      s = wave_sum(s);
      if (lane == 0) L.rms[c] = sqrt(s / (double)n);
    }
    __syncthreads();
    for (int r = tid; r < n; r += blockDim.x)
      for (int c = 0; c < kb; ++c) Xs[(size_t)r * kb + c] /= L.rms[c];
    __syncthreads();
    // ---- ||X||_inf
    double part = 0.0;
    for (int r = tid; r < n; r += blockDim.x) {
      double s = 0.0;
      for (int c = 0; c < kb; ++c) s += fabs(Xs[(size_t)r * kb + c]);
      part = fmax(part, s);
    }
    const double xnorm = block_max(part, L.red);
    // ---- column means (long double in the reference) for the constraint
    for (int c = wid; c < kb; c += nw) {
      dd s{0.0, 0.0};
      for (int r = lane; r < n; r += KPC_WAVE) s = dd_add(s, dd{Xs[(size_t)r * kb + c], 0.0});
      s = wave_sum_dd(s);
      if (lane == 0) {
        dd m = dd_div_int(s, (double)n);
        L.cc[c] = m.hi + m.lo;
      }
    }
    __syncthreads();
    if (tid == 0) {
      // penalty P = Z' D Z with the reflectors, zero null space, RMS scaling
      double* P = Xs + (size_t)n * kb;  // kb x kb row-major in scratch
      for (int i = 0; i < kb * kb; ++i) P[i] = 0.0;
      for (int i = 0; i < kb; ++i) P[i * kb + i] = L.cm[i];
      double prod[kMaxBasis];
      for (int t = 0; t < M; ++t) {  // hq_post
        const double* ht = L.z + t * kb;
        for (int r = 0; r < kb; ++r) {
          double v = 0.0;
          for (int c = 0; c < kb; ++c) v += P[r * kb + c] * ht[c];
          prod[r] = v;
        }
        for (int r = 0; r < kb; ++r)
          for (int c = 0; c < kb; ++c) P[r * kb + c] -= prod[r] * ht[c];
      }
      for (int t = 0; t < M; ++t) {  // hq_pre_transpose
        const double* ht = L.z + t * kb;
        for (int c = 0; c < kb; ++c) {
          double v = 0.0;
          for (int r = 0; r < kb; ++r) v += P[r * kb + c] * ht[r];
          prod[c] = v;
        }
        for (int r = 0; r < kb; ++r)
          for (int c = 0; c < kb; ++c) P[r * kb + c] -= prod[c] * ht[r];
      }
      for (int r = 0; r < kb; ++r)
        for (int c = kb - M; c < kb; ++c) { P[r * kb + c] = 0.0; P[c * kb + r] = 0.0; }
      for (int c = 0; c < kb; ++c) {
        for (int i = 0; i < kb; ++i) P[c * kb + i] /= L.rms[c];
        for (int i = 0; i < kb; ++i) P[i * kb + c] /= L.rms[c];
      }
      double* SP = P + kb * kb;
      for (int r = 0; r < kb; ++r)
        for (int c = 0; c < kb; ++c) SP[r * kb + c] = (P[r * kb + c] + P[c * kb + r]) / 2.0;
      double pnorm = 0.0;
      for (int c = 0; c < kb; ++c) {
        double s = 0.0;
        for (int r = 0; r < kb; ++r) s += fabs(SP[r * kb + c]);
        if (s > pnorm) pnorm = s;
      }
      const double pscale = pnorm / (xnorm * xnorm);
      for (int i = 0; i < kb * kb; ++i) SP[i] /= pscale;
      // LINPACK single-column Householder on the column means
      double* cmv = L.cc;
      double nrm = 0.0;
      for (int i = 0; i < kb; ++i) nrm += cmv[i] * cmv[i];
      nrm = sqrt(nrm);
      double qraux = 0.0;
      if (nrm != 0.0) {
        if (cmv[0] != 0.0) nrm = copysign(nrm, cmv[0]);
        double rr = 1.0 / nrm;
        for (int i = 0; i < kb; ++i) cmv[i] *= rr;
        cmv[0] = 1.0 + cmv[0];
        qraux = cmv[0];
        cmv[0] = -nrm;
      }
      L.dctl[1] = qraux;
      // absorbed penalty: Q' S Q without first row/col
      double* left = SP + kb * kb;  // column-major kb x kb
      double yv[kMaxBasis];
      for (int c = 0; c < kb; ++c) {
        for (int r = 0; r < kb; ++r) yv[r] = SP[r * kb + c];
        if (qraux != 0.0) {
          double t = qraux * yv[0];
          for (int i = 1; i < kb; ++i) t += cmv[i] * yv[i];
          t = -t / qraux;
          yv[0] += t * qraux;
          for (int i = 1; i < kb; ++i) yv[i] += t * cmv[i];
        }
        for (int r = 0; r < kb; ++r) left[c * kb + r] = yv[r];
      }
      double* So = out_S + (size_t)slot * 29 * 29;
      for (int fr = 0; fr < kb - 1; ++fr) {
        for (int oc = 0; oc < kb; ++oc) yv[oc] = left[oc * kb + fr + 1];
        if (qraux != 0.0) {
          double t = qraux * yv[0];
          for (int i = 1; i < kb; ++i) t += cmv[i] * yv[i];
          t = -t / qraux;
          yv[0] += t * qraux;
          for (int i = 1; i < kb; ++i) yv[i] += t * cmv[i];
        }
        for (int c = 0; c < kb - 1; ++c) So[c * (kb - 1) + fr] = yv[c + 1];
      }
      info[0] = 0;
      info[1] = nu;
      info[2] = completed;
      info[3] = kb - M;
    }
    __syncthreads();
    // ---- absorbed X rows -> column-major n x (kb-1)
    {
      const double qraux = L.dctl[1];
      const double* cmv = L.cc;
      double* Xo = out_X + (size_t)slot * n * 29;
      for (int r = tid; r < n; r += blockDim.x) {
        double yv[kMaxBasis];
        for (int c = 0; c < kb; ++c) yv[c] = Xs[(size_t)r * kb + c];
        if (qraux != 0.0) {
          double t = qraux * yv[0];
          for (int i = 1; i < kb; ++i) t += cmv[i] * yv[i];
          t = -t / qraux;
          yv[0] += t * qraux;
          for (int i = 1; i < kb; ++i) yv[i] += t * cmv[i];
        }
        for (int c = 0; c < kb - 1; ++c) Xo[(size_t)c * n + r] = yv[c + 1];
      }
    }
    __syncthreads();
  }
}

long long smooth_scratch_doubles(int n) {
  // E + Q basis + QL vectors + U + X (row-major) + penalty work
  long long nu = n;
  long long s = nu * nu + nu * (nu + 1) + nu * nu + nu * kMaxBasis + (long long)n * kMaxBasis +
                3LL * kMaxBasis * kMaxBasis + 64;
  if (s < 2LL * n) s = 2LL * n;
  return (s + 31) / 32 * 32;
}

size_t smooth_lds_bytes(int n) {
  int np2 = 1;
  while (np2 < n) np2 <<= 1;
  return smooth_lds_size(np2);
}

hipError_t launch_smooth_setup(const double* data, int n, int p, const int* items, int count,
                               double* out_X, double* out_S, int* out_info, double* scratch,
                               int grid, hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  const size_t lds = smooth_lds_bytes(n);
  hipError_t err = hipFuncSetAttribute((const void*)smooth_setup_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(smooth_setup_kernel, dim3(grid), dim3(kSmoothThreads), lds,
                     stream, data, n, p, items, count, out_X, out_S, out_info, scratch,
                     smooth_scratch_doubles(n));
  return hipGetLastError();
}

}  // namespace kpc
