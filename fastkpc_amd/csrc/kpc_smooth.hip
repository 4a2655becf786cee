// Thin-plate regression smooth construction (mgcv 1.9-1 `tp`, k = 10 for one
// covariate, 30 for two) on gfx950: one workgroup builds one smooth.
//
// Reference semantics: fastkpc/src/full_cuda_ci_native_setup.cpp
//   :96-107  shift = R mean of each covariate (long double + correction)
//   :114-152 eta constants, :154-198 null space / polynomial powers
//   :200-239 lexicographically sorted unique locations
//   :266-301 eigenpairs of the Lanczos tridiagonal, descending order
//   :303-494 Rlanczos: LCG start vector, Lanczos with two reorthogonalisation
//            passes, Ritz check every max(10, k/2) (capped at n/10) steps
//            once the step index reaches k, eps^0.7 * ||T|| stop rule with the
//            lower-end index quirk at :439, Ritz vector assembly
//   :496-594 QT Householders, hq_post, hq_pre_transpose
//   :621-688 sum-to-zero constraint (LINPACK dqrdc2/dqrqty, one reflector)
//   :690-858 X = [U D Z | T], penalty Z' D Z, RMS column scaling,
//            S <- S / (||S||_1 / ||X||_inf^2)
//
// Device design:
//  * E_ij = eta(|x_i - x_j|) is evaluated once per smooth, not once per
//    Lanczos step: the upper triangle of E in 64 x 64 tiles (T(T+1)/2 tiles,
//    T = ceil(nu / 64), zero past nu) goes to the workgroup's slab, each tile
//    stored by diagonals (entry t * 64 + l is E[64 I + l][64 J + (l + t) mod
//    64]) so a wave reads it with 64 coalesced loads.  A step then streams
//    the tiles once: lane l adds E_ij q_j to its row, and an off-diagonal tile
//    serves its transpose as well through a column accumulator that rotates
//    one lane per diagonal (wave_rol:1), so lane l meets column (l + t) mod
//    64 at diagonal t.  Tile partials are summed per row in tile order
//    (deterministic).  A step costs 8 bytes and ~4 instructions per stored
//    entry instead of ~48 FP64 instructions per entry of E.
//  * The eta of d = 2 uses tp_log (atanh series, within 2 ulp) instead of
//    the library logarithm, whose double-double compensation triples the
//    cost of the product.
//  * The reference's two modified Gram-Schmidt passes after the three-term
//    recurrence become classical Gram-Schmidt passes, the second one only
//    when the first leaves less than half the squared norm (DGKS); both
//    schemes are orthogonal to working precision.  A wave forms eight
//    coefficients at once with coalesced, unconditional basis reads.
//  * A Ritz check needs the k largest and k smallest eigenvalues of the
//    Lanczos tridiagonal and the last components of the k leading
//    eigenvectors: Sturm multisection and twisted factorisations give them
//    in parallel, and the Ritz vectors of the selected pairs come from
//    twisted factorisations as well (re-orthogonalised within eigenvalue
//    clusters).  Eigenvector signs are immaterial: flipping a Ritz vector
//    flips the matching column of the basis and rows/columns of the penalty,
//    which leaves the fitted smooth unchanged.
//  * The penalty transforms run in LDS with one thread per row or column;
//    each entry is formed with the reference's operation order.
#include "kpc_common.h"
#include "kpc_device.h"
#include "kpc_tridiag.h"

namespace kpc {

constexpr int kSmoothThreads = 512;
constexpr int kMaxBasis = 32;
constexpr int kSmoothWork = 4096;  // doubles of LDS work area

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

// Natural logarithm of a finite x > 0 by argument reduction and the atanh
// series: x = 2^k m with m in [sqrt(1/2), sqrt(2)), s = (m - 1) / (m + 1),
// log(m) = 2 atanh(s) = 2 s sum_{i=0..11} s^(2i) / (2i + 1) (|s| < 0.1716, so
// the omitted terms are below 2^-56 relative), log 2 split into a 32-bit head
// (k log2_head exact) and a tail.  Within 2 ulp of the correctly rounded
// logarithm; a third of the instructions of the library routine.  x <= 0
// gives a meaningless finite value: callers mask it.
__device__ __forceinline__ double tp_log(double x) {
  double m = __builtin_amdgcn_frexp_mant(x);
  int k = __builtin_amdgcn_frexp_exp(x);
  if (m < 0.7071067811865476) {
    m = m + m;
    k -= 1;
  }
  const double f = m - 1.0;
  const double den = m + 1.0;
  double r = __builtin_amdgcn_rcp(den);
  r = fma(r, fma(-den, r, 1.0), r);
  r = fma(r, fma(-den, r, 1.0), r);
  double s = f * r;
  s = fma(r, fma(-den, s, f), s);
  const double z = s * s;
  double P = 0.043478260869565216;   // 1/23
  P = fma(P, z, 0.047619047619047616);  // 1/21
  P = fma(P, z, 0.05263157894736842);   // 1/19
  P = fma(P, z, 0.058823529411764705);  // 1/17
  P = fma(P, z, 0.06666666666666667);   // 1/15
  P = fma(P, z, 0.07692307692307693);   // 1/13
  P = fma(P, z, 0.09090909090909091);   // 1/11
  P = fma(P, z, 0.1111111111111111);    // 1/9
  P = fma(P, z, 0.14285714285714285);   // 1/7
  P = fma(P, z, 0.2);                   // 1/5
  P = fma(P, z, 0.3333333333333333);    // 1/3
  const double lm = 2.0 * fma(s * z, P, s);
  const double kd = (double)k;
  return fma(kd, 0.6931471803691238, fma(kd, 1.9082149292705877e-10, lm));
}

// Order-2 thin-plate eta from the coordinate differences: d = 1 gives
// eta0 r^2 |r| (sqrt(fl(dx^2)) = |dx| exactly under correct rounding), d = 2
// gives eta0 (log(r^2) / 2) r^2 evaluated as (eta0 / 2) log(r^2) r^2 (the
// halving is exact); zero at r = 0.  eta = eta0 (d = 1) or eta0 / 2 (d = 2).
template <int D>
__device__ __forceinline__ double tp_eta(double dx, double dy, double eta) {
  if (D == 2) {
    const double r2 = dx * dx + dy * dy;
    const double v = (eta * tp_log(r2)) * r2;
    return r2 > 0.0 ? v : 0.0;
  }
  const double r2 = dx * dx;
  const double v = (eta * r2) * fabs(dx);
  return r2 > 0.0 ? v : 0.0;
}

// Workgroup LDS, carved from dynamic shared memory with arrays of np2 (the
// next power of two >= n) entries.
struct SmoothLds {
  double* ka;    // sort key 0 -> unique coordinate 0
  double* kb;    // sort key 1 -> unique coordinate 1
  double* qv;    // current Lanczos vector
  double* z;     // Lanczos residual (unique ids while sorting)
  double* al;    // Lanczos alpha
  double* be;    // Lanczos beta
  double* e2;    // beta^2 for the Sturm counts
  double* w;     // kSmoothWork: tile partials of E q, CGS coefficients,
                 // twisted pivots, penalty work
  double* red;   // 16
  double* dctl;  // 8
  double* zq;    // 3 * kMaxBasis  T'U rows
  double* hz;    // 3 * kMaxBasis  QT reflectors
  double* cm;    // kMaxBasis      Ritz values of the basis
  double* rms;   // kMaxBasis
  double* clo;   // 2 * kMaxBasis  multisection brackets
  double* chi;   // 2 * kMaxBasis
  double* cval;  // 2 * kMaxBasis  candidate eigenvalues (ascending slots)
  double* errs;  // kMaxBasis      Ritz residual bounds of the leading pairs
  int* ki;       // sort index -> original row -> unique id
  int* ctl;      // 16
};

__host__ __device__ inline size_t smooth_lds_size(int np2) {
  return (size_t)np2 * (7 * sizeof(double) + sizeof(int)) +
         (size_t)(kSmoothWork + 16 + 8 + 6 * kMaxBasis + 2 * kMaxBasis + 6 * kMaxBasis +
                  kMaxBasis) * sizeof(double) +
         16 * sizeof(int);
}

// Dynamic LDS of a set-up (one workgroup per CU) is capped at this: the
// fixed areas above, then a cache of the first Lanczos basis columns (column
// stride n, whole Gram-Schmidt blocks of eight; none when under eight fit).
constexpr size_t kSmoothLdsBudget = 163840 - 1024;

__host__ __device__ inline int smooth_cache_cols(int n) {
  int np2 = 64;
  while (np2 < n) np2 <<= 1;
  const size_t base = smooth_lds_size(np2);
  if (base >= kSmoothLdsBudget) return 0;
  const int c = (int)((kSmoothLdsBudget - base) / (sizeof(double) * (size_t)n));
  return (c < 64 ? c : 64) & ~7;
}

__device__ inline SmoothLds carve_smooth_lds(double* base, int np2) {
  SmoothLds L;
  double* d = base;
  L.ka = d; d += np2;
  L.kb = d; d += np2;
  L.qv = d; d += np2;
  L.z = d; d += np2;
  L.al = d; d += np2;
  L.be = d; d += np2;
  L.e2 = d; d += np2;
  L.w = d; d += kSmoothWork;
  L.red = d; d += 16;
  L.dctl = d; d += 8;
  L.zq = d; d += 3 * kMaxBasis;
  L.hz = d; d += 3 * kMaxBasis;
  L.cm = d; d += kMaxBasis;
  L.rms = d; d += kMaxBasis;
  L.clo = d; d += 2 * kMaxBasis;
  L.chi = d; d += 2 * kMaxBasis;
  L.cval = d; d += 2 * kMaxBasis;
  L.errs = d; d += kMaxBasis;
  int* ip = reinterpret_cast<int*>(d);
  L.ki = ip; ip += np2;
  L.ctl = ip;
  return L;
}

// Tiles of the upper triangle of E: T = ceil(nu / 64) row blocks, tile s =
// (I, J) with I <= J in row-major order, 4096 doubles each.
__host__ __device__ inline int smooth_tile_dim(int nu) { return (nu + 63) / 64; }
__host__ __device__ inline int smooth_tile_count(int nu) {
  const int T = smooth_tile_dim(nu);
  return T * (T + 1) / 2;
}
__device__ __forceinline__ int tile_index(int I, int J, int T) {
  return I * T - I * (I - 1) / 2 + (J - I);
}
__device__ __forceinline__ void tile_decode(int s, int T, int& I, int& J) {
  I = 0;
  while (s >= T - I) {
    s -= T - I;
    ++I;
  }
  J = I + s;
}

// Value of lane (lane + 1) mod 64 (DPP wave_rol:1; every lane is active).
__device__ __forceinline__ double wave_rol1(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x134, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x134, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// E tiles into Et, stored by diagonals: Et[s][t][l] = E[64 I + l][64 J +
// (l + t) mod 64], zero outside nu x nu.  Each wave builds whole tiles.
template <int D>
__device__ void tp_build_tiles(const SmoothLds& L, int nu, double eta, double* __restrict__ Et) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int wid = threadIdx.x / KPC_WAVE;
  const int nw = blockDim.x / KPC_WAVE;
  const int T = smooth_tile_dim(nu), ntile = smooth_tile_count(nu);
  for (int s = wid; s < ntile; s += nw) {
    int I, J;
    tile_decode(s, T, I, J);
    const int i = 64 * I + lane;
    const bool iv = i < nu;
    const double xi = iv ? L.ka[i] : 0.0;
    const double yi = iv && D == 2 ? L.kb[i] : 0.0;
    double* out = Et + (size_t)s * 4096 + lane;
    for (int t = 0; t < 64; ++t) {
      const int j = 64 * J + ((lane + t) & 63);
      double e = 0.0;
      if (iv && j < nu) e = tp_eta<D>(xi - L.ka[j], D == 2 ? yi - L.kb[j] : 0.0, eta);
      out[64 * t] = e;
    }
  }
}

// Tile partials of E q: part[s][0][l] = sum_t Et[s][t][l] q[64 J + (l+t) mod
// 64] (rows 64 I + l), and for I < J part[s][1][l] = sum_t Et[s][t][(l - t)
// mod 64] q[64 I + (l - t) mod 64] (rows 64 J + l, the transpose) through the
// rotating column accumulator.  q must be zero from nu to 64 T.
__device__ void tp_tile_product(const double* __restrict__ Et, const double* q, int nu,
                                double* part) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int wid = threadIdx.x / KPC_WAVE;
  const int nw = blockDim.x / KPC_WAVE;
  const int T = smooth_tile_dim(nu), ntile = smooth_tile_count(nu);
  for (int s = wid; s < ntile; s += nw) {
    int I, J;
    tile_decode(s, T, I, J);
    const double* __restrict__ tile = Et + (size_t)s * 4096 + lane;
    const double* qJ = q + 64 * J;
    const double qi = q[64 * I + lane];
    double row = 0.0, col = 0.0;
    // the tile's diagonals in two batches of 32 loads, all issued before the
    // first is used
    for (int t0 = 0; t0 < 64; t0 += 32) {
      double e[32];
#pragma unroll
      for (int u = 0; u < 32; ++u) e[u] = tile[64 * (t0 + u)];
      if (I == J) {
#pragma unroll
        for (int u = 0; u < 32; ++u) row = fma(e[u], qJ[(lane + t0 + u) & 63], row);
      } else {
#pragma unroll
        for (int u = 0; u < 32; ++u) {
          row = fma(e[u], qJ[(lane + t0 + u) & 63], row);
          if (t0 + u > 0) col = wave_rol1(col);
          col = fma(e[u], qi, col);
        }
      }
    }
    if (I != J) col = wave_rol1(col);
    part[(size_t)s * 128 + lane] = row;
    part[(size_t)s * 128 + 64 + lane] = col;
  }
}

// Gram-Schmidt coefficient partial sums of z against basis columns b0 ..
// b0 + nb - 1 (lanes over rows; columns past nb repeat the last one), the
// columns read from qbase: the slab or its LDS copy.
__device__ __forceinline__ void smooth_block_coeffs(const double* qbase, const double* z, int nu,
                                                    int b0, int nb, int lane, double g[8]) {
  const double* qrow[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) qrow[k] = qbase + (size_t)(b0 + (k < nb ? k : nb - 1)) * nu;
#pragma unroll 2
  for (int i = lane; i < nu; i += KPC_WAVE) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = qrow[k][i];
    const double zi = z[i];
#pragma unroll
    for (int k = 0; k < 8; ++k) g[k] += zi * v[k];
  }
}

// Row i's sum over whole blocks b0 <= b < b1 of eight basis columns read from
// qbase, accumulated into sa[b mod 8].
__device__ __forceinline__ void smooth_block_sums(const double* qbase, const double* hc, int nu,
                                                  int i, int b0, int b1, double sa[8]) {
#pragma unroll 2
  for (int b = b0; b + 8 <= b1; b += 8) {
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = qbase[(size_t)(b + k) * nu + i];
#pragma unroll
    for (int k = 0; k < 8; ++k) sa[k] += hc[b + k] * v[k];
  }
}

__global__ __launch_bounds__(kSmoothThreads) void smooth_setup_kernel(
    const double* __restrict__ data, int n, int p, const int* __restrict__ items, int count,
    double* __restrict__ out_X, double* __restrict__ out_S, int* __restrict__ out_info,
    double* __restrict__ scratch, long long scratch_stride) {
  extern __shared__ double lds_raw[];
  constexpr double kEps = 2.220446049250313e-16;
  int np2 = 64;  // at least one tile row of q (zero past nu)
  while (np2 < n) np2 <<= 1;
  SmoothLds L = carve_smooth_lds(lds_raw, np2);
  // basis columns 0 .. ncache - 1 of the current set-up, column stride nu:
  // the Gram-Schmidt passes read them here instead of the slab (same values)
  double* qc = lds_raw + smooth_lds_size(np2) / sizeof(double);
  const int ncache0 = smooth_cache_cols(n);
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
    double* Qb = ws;                                 // nu x (nu+1) Lanczos basis
    double* Zs = Qb + (size_t)nu * (nu + 1);         // kMaxBasis x nu Ritz coefficients
    double* U = Zs + (size_t)kMaxBasis * nu;         // nu x kb Ritz vectors
    double* Xs = U + (size_t)nu * kMaxBasis;         // n x kb (row-major) before absorption
    double* Et = Xs + (size_t)n * kMaxBasis;         // E tiles (4096 each)
    const int ntile = smooth_tile_count(nu);
    // tile partials of E q and the Gram-Schmidt coefficients: in the LDS work
    // area when both fit, else partials in the slab after the tiles
    const bool part_lds = (long long)ntile * 128 + nu <= kSmoothWork;
    double* tpart = part_lds ? L.w : Et + (size_t)ntile * 4096;
    double* hc = part_lds ? L.w + (size_t)ntile * 128 : L.w;  // Gram-Schmidt coefficients (nu)
    const double eta0 = eta_constant(order, d);
    const double eta = d == 2 ? 0.5 * eta0 : eta0;
    if (d == 2)
      tp_build_tiles<2>(L, nu, eta, Et);
    else
      tp_build_tiles<1>(L, nu, eta, Et);
    for (int r = nu + tid; r < 64 * smooth_tile_dim(nu); r += blockDim.x) L.qv[r] = 0.0;
    // ---- Lanczos start vector (mgcv LCG)
    if (tid == 0) {
      unsigned long st = 1;
      double nrm = 0.0;
      for (int r = 0; r < nu; ++r) {
        st = (st * 106UL + 1283UL) % 6075UL;
        double vv = (double)st / 6075.0 - 0.5;
        L.z[r] = vv;
        nrm += vv * vv;
      }
      L.dctl[0] = sqrt(nrm);
    }
    __syncthreads();
    const int ncache = ncache0 < nu + 1 ? ncache0 : (nu + 1) & ~7;
    for (int r = tid; r < nu; r += blockDim.x) {
      const double v = L.z[r] / L.dctl[0];
      L.qv[r] = v;
      Qb[r] = v;
      if (ncache > 0) qc[r] = v;
    }
    __syncthreads();
    const double tol = pow(kEps, 0.7);
    int check = kb / 2;
    if (check < 10) check = 10;
    int tenth = (int)floor(nu / 10.0);
    if (tenth < 1) tenth = 1;
    if (tenth < check) check = tenth;
    const int T = smooth_tile_dim(nu);
    int upper = kb, lower = 0, completed = nu;
    for (int it = 0; it < nu; ++it) {
      __syncthreads();  // q complete (and, at it = 0, the tiles)
      tp_tile_product(Et, L.qv, nu, tpart);
      __syncthreads();
      // z = E q_it (tile partials of row block R in tile order: the transposed
      // parts of tiles (J, R), J < R, then the row parts of tiles (R, J)),
      // alpha, three-term recurrence
      double part = 0.0;
      for (int r = tid; r < nu; r += blockDim.x) {
        const int R = r >> 6, l = r & 63;
        double sz = 0.0;
        for (int J = 0; J < R; ++J) sz += tpart[(size_t)tile_index(J, R, T) * 128 + 64 + l];
        for (int J = R; J < T; ++J) sz += tpart[(size_t)tile_index(R, J, T) * 128 + l];
        L.z[r] = sz;
        part += L.qv[r] * sz;
      }
      const double proj = block_sum(part, L.red);
      if (tid == 0) L.al[it] = proj;
      const double prior = it > 0 ? L.be[it - 1] : 0.0;
      const double* qp = Qb + (size_t)(it > 0 ? it - 1 : 0) * nu;
      part = 0.0;
      for (int r = tid; r < nu; r += blockDim.x) {
        const double v = L.z[r] - (proj * L.qv[r] + prior * qp[r]);
        L.z[r] = v;
        part += v * v;
      }
      double zn2 = block_sum(part, L.red);
      // Reorthogonalisation (it > 0; at it = 0 the reference only removes the
      // alpha component): classical Gram-Schmidt against Q[0..it], repeated
      // once more when a pass leaves less than half the squared norm it
      // started from (DGKS).  A wave forms eight coefficients at once (rows
      // past the last basis vector are clamped to it, so every load is
      // unconditional and all of them can be in flight together).
      for (int pass = 0; pass < 2 && it > 0; ++pass) {
        for (int b0 = wid * 8; b0 <= it; b0 += nw * 8) {
          double g[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
          const int nb = min(8, it + 1 - b0);
          if (b0 < ncache)
            smooth_block_coeffs(qc, L.z, nu, b0, nb, lane, g);
          else
            smooth_block_coeffs(Qb, L.z, nu, b0, nb, lane, g);
#pragma unroll
          for (int k = 0; k < 8; ++k) g[k] = wave_sum(g[k]);
          if (lane < nb) {
            double gv = g[0];
#pragma unroll
            for (int k = 1; k < 8; ++k) gv = lane == k ? g[k] : gv;
            hc[b0 + lane] = gv;
          }
        }
        __syncthreads();
        part = 0.0;
        for (int i = tid; i < nu; i += blockDim.x) {
          double sa[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
          const int bc = min((it + 1) & ~7, ncache);  // whole blocks held in LDS
          smooth_block_sums(qc, hc, nu, i, 0, bc, sa);
          int b = bc;
#pragma unroll 2
          for (; b + 8 <= it + 1; b += 8) {
            double v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = Qb[(size_t)(b + k) * nu + i];
#pragma unroll
            for (int k = 0; k < 8; ++k) sa[k] += hc[b + k] * v[k];
          }
          for (; b <= it; ++b)
            sa[0] += hc[b] * (b < ncache ? qc[(size_t)b * nu + i] : Qb[(size_t)b * nu + i]);
          const double zi =
              L.z[i] - (((sa[0] + sa[1]) + (sa[2] + sa[3])) + ((sa[4] + sa[5]) + (sa[6] + sa[7])));
          L.z[i] = zi;
          part += zi * zi;
        }
        const double after = block_sum(part, L.red);
        const bool again = after < 0.5 * zn2;
        zn2 = after;
        if (!again) break;
      }
      const double rn = sqrt(zn2);
      if (tid == 0) L.be[it] = rn;
      if (it < nu - 1) {
        double* qn = Qb + (size_t)(it + 1) * nu;
        for (int r = tid; r < nu; r += blockDim.x) {
          const double v = L.z[r] / rn;
          L.qv[r] = v;
          qn[r] = v;
          if (it + 1 < ncache) qc[(size_t)(it + 1) * nu + r] = v;
        }
      }
      __syncthreads();
      if (!((it >= kb && it % check == 0) || it == nu - 1)) continue;
      // ---- Ritz check on the tridiagonal T (size x size)
      const int size = it + 1;
      double gl = 1e308, gu = -1e308, emx = 0.0;
      for (int i = tid; i < size; i += blockDim.x) {
        const double rad = (i > 0 ? fabs(L.be[i - 1]) : 0.0) + (i < size - 1 ? fabs(L.be[i]) : 0.0);
        gl = fmin(gl, L.al[i] - rad);
        gu = fmax(gu, L.al[i] + rad);
        if (i < size - 1) {
          const double b2 = L.be[i] * L.be[i];
          L.e2[i] = b2;
          emx = fmax(emx, b2);
        }
      }
      gl = -block_max(-gl, L.red);
      gu = block_max(gu, L.red);
      emx = block_max(emx, L.red);
      const double pivmin = 2.2250738585072014e-308 * fmax(1.0, emx);
      const double span = gu - gl;
      const double lo0 = gl - 2.0 * kEps * span - pivmin;
      const double hi0 = gu + 2.0 * kEps * span + pivmin;
      const int kk = kb < size ? kb : size;
      const int nlow = kk;
      const int hs = (size - kk) > nlow ? (size - kk) : nlow;
      const int ncand = nlow + (size - hs);
      lz_multisect(L.al, L.e2, size, pivmin, ncand, nlow, hs, 2.0 * kEps,
                   2.0 * kEps * fmax(fabs(lo0), fabs(hi0)), true, lo0, hi0, L.clo, L.chi, L.cval);
      // eigenvalue with descending index j
      auto desc_val = [&](int j) {
        const int a = size - 1 - j;
        return L.cval[a < nlow ? a : nlow + (a - hs)];
      };
      const int per = 2 * size;
      const int B = kSmoothWork / per > 0 ? kSmoothWork / per : 1;
      for (int j0 = 0; j0 < kk; j0 += B) {
        const int j = j0 + tid;
        if (tid < B && j < kk) {
          double* wsl = L.w + (size_t)tid * per;
          const double last = lz_twisted(L.al, L.be, size, desc_val(j), pivmin, wsl, wsl + size,
                                         nullptr);
          L.errs[j] = fabs(rn * last);
        }
        __syncthreads();
      }
      if (tid == 0) {
        double tnorm = fabs(desc_val(0));
        if (fabs(desc_val(it)) > tnorm) tnorm = fabs(desc_val(it));
        int conv = 0, pos = 0, neg = 0;
        if (it >= kb) {
          const double maxerr = tnorm * tol;
          conv = 1;
          while (pos + neg < kb) {
            if (fabs(desc_val(pos)) >= fabs(desc_val(it - neg))) {
              if (L.errs[pos] > maxerr) { conv = 0; break; }
              ++pos;
            } else {
              if (L.errs[neg] > maxerr) { conv = 0; break; }  // reference index quirk
              ++neg;
            }
          }
        }
        L.ctl[6] = conv;
        L.ctl[7] = pos;
        L.ctl[8] = neg;
      }
      __syncthreads();
      if (!L.ctl[6] && it < nu - 1) continue;
      if (L.ctl[6]) {
        upper = L.ctl[7];
        lower = L.ctl[8];
      }
      completed = size;
      // ---- coefficient vectors of the selected Ritz pairs into Zs (K x size)
      const int K = upper + lower;
      for (int e0 = 0; e0 < K; e0 += B) {
        const int e = e0 + tid;
        if (tid < B && e < K) {
          const int src = e < upper ? e : size - (upper + lower - e);
          const double th = desc_val(src);
          double* wsl = L.w + (size_t)tid * per;
          lz_twisted(L.al, L.be, size, th, pivmin, wsl, wsl + size, Zs + (size_t)e * size);
          L.cm[e] = th;
        }
        __syncthreads();
      }
      // re-orthogonalise within eigenvalue clusters (relative gap < 1e-3)
      if (wid == 0) {
        for (int e = 1; e < K; ++e) {
          const double le = L.cm[e];
          double* ze = Zs + (size_t)e * size;
          bool touched = false;
          for (int f = 0; f < e; ++f) {
            const double lf = L.cm[f];
            if (fabs(le - lf) > 1e-3 * fmax(fabs(le), fabs(lf))) continue;
            const double* zf = Zs + (size_t)f * size;
            double s = 0.0;
            for (int i = lane; i < size; i += KPC_WAVE) s += zf[i] * ze[i];
            s = wave_sum(s);
            for (int i = lane; i < size; i += KPC_WAVE) ze[i] -= s * zf[i];
            touched = true;
          }
          if (touched) {
            double s = 0.0;
            for (int i = lane; i < size; i += KPC_WAVE) s += ze[i] * ze[i];
            s = 1.0 / sqrt(wave_sum(s));
            for (int i = lane; i < size; i += KPC_WAVE) ze[i] *= s;
          }
        }
      }
      __syncthreads();
      break;
    }
    // ---- Ritz vectors U = Q Zs (nu x K) on matrix cores (each entry's sum
    // over the basis in order, as a chain of FMAs)
    ritz_assemble_mfma(Qb, Zs, completed, nu, completed, upper + lower, nullptr, U);
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
      // QT householders: zq (input, row-major) -> stored reflectors in hz[]
      double* in = L.zq;
      double* st = L.hz;  // M x kb
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
        const double* ht = L.hz + t * kb;
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
    // ---- penalty work in LDS: P, SP (row-major kb x kb), left (column-major),
    // column means; prod in the (now idle) multisection brackets
    double* P = L.w;
    double* SP = P + kMaxBasis * kMaxBasis;
    double* left = SP + kMaxBasis * kMaxBasis;
    double* cmv = left + kMaxBasis * kMaxBasis;
    double* prod = L.clo;
    const int kk2 = kb * kb;
    // ---- column means (long double in the reference) for the constraint
    for (int c = wid; c < kb; c += nw) {
      dd s{0.0, 0.0};
      for (int r = lane; r < n; r += KPC_WAVE) s = dd_add(s, dd{Xs[(size_t)r * kb + c], 0.0});
      s = wave_sum_dd(s);
      if (lane == 0) {
        dd m = dd_div_int(s, (double)n);
        cmv[c] = m.hi + m.lo;
      }
    }
    // ---- penalty P = Z' D Z: diag(D), hq_post then hq_pre_transpose
    for (int i = tid; i < kk2; i += blockDim.x) P[i] = (i / kb == i % kb) ? L.cm[i / kb] : 0.0;
    __syncthreads();
    for (int t = 0; t < M; ++t) {  // P <- P (I - h h')
      const double* ht = L.hz + t * kb;
      for (int r = tid; r < kb; r += blockDim.x) {
        double v = 0.0;
        for (int c = 0; c < kb; ++c) v += P[r * kb + c] * ht[c];
        prod[r] = v;
      }
      __syncthreads();
      for (int i = tid; i < kk2; i += blockDim.x) P[i] -= prod[i / kb] * ht[i % kb];
      __syncthreads();
    }
    for (int t = 0; t < M; ++t) {  // P <- (I - h h') P
      const double* ht = L.hz + t * kb;
      for (int c = tid; c < kb; c += blockDim.x) {
        double v = 0.0;
        for (int r = 0; r < kb; ++r) v += P[r * kb + c] * ht[r];
        prod[c] = v;
      }
      __syncthreads();
      for (int i = tid; i < kk2; i += blockDim.x) P[i] -= prod[i % kb] * ht[i / kb];
      __syncthreads();
    }
    // zero null-space rows/columns; RMS scaling (row c then column c for
    // ascending c: an entry is divided by the smaller index's rms first)
    for (int i = tid; i < kk2; i += blockDim.x) {
      const int r = i / kb, c = i % kb;
      double v = (r >= kb - M || c >= kb - M) ? 0.0 : P[i];
      const int a = r < c ? r : c, b = r < c ? c : r;
      v /= L.rms[a];
      v /= L.rms[b];
      P[i] = v;
    }
    __syncthreads();
    for (int i = tid; i < kk2; i += blockDim.x) {
      const int r = i / kb, c = i % kb;
      SP[i] = (P[r * kb + c] + P[c * kb + r]) / 2.0;
    }
    __syncthreads();
    double cpart = 0.0;
    for (int c = tid; c < kb; c += blockDim.x) {
      double s = 0.0;
      for (int r = 0; r < kb; ++r) s += fabs(SP[r * kb + c]);
      cpart = s;
    }
    const double pnorm = block_max(cpart, L.red);
    const double pscale = pnorm / (xnorm * xnorm);
    for (int i = tid; i < kk2; i += blockDim.x) SP[i] /= pscale;
    if (tid == 0) {
      // LINPACK single-column Householder on the column means
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
    }
    __syncthreads();
    const double qraux = L.dctl[1];
    // absorbed penalty: Q' S Q without the first row/column; left = Q' SP
    for (int c = tid; c < kb; c += blockDim.x) {
      double* yv = left + c * kb;
      for (int r = 0; r < kb; ++r) yv[r] = SP[r * kb + c];
      if (qraux != 0.0) {
        double t = qraux * yv[0];
        for (int i = 1; i < kb; ++i) t += cmv[i] * yv[i];
        t = -t / qraux;
        yv[0] += t * qraux;
        for (int i = 1; i < kb; ++i) yv[i] += t * cmv[i];
      }
    }
    __syncthreads();
    {
      double* So = out_S + (size_t)slot * 29 * 29;
      for (int fr = tid; fr < kb - 1; fr += blockDim.x) {
        double t = 0.0;
        if (qraux != 0.0) {
          t = qraux * left[fr + 1];
          for (int i = 1; i < kb; ++i) t += cmv[i] * left[i * kb + fr + 1];
          t = -t / qraux;
        }
        for (int c = 0; c < kb - 1; ++c) {
          double v = left[(c + 1) * kb + fr + 1];
          if (qraux != 0.0) v += t * cmv[c + 1];
          So[c * (kb - 1) + fr] = v;
        }
      }
    }
    if (tid == 0) {
      info[0] = 0;
      info[1] = nu;
      info[2] = completed;
      info[3] = kb - M;
    }
    // ---- absorbed X rows -> column-major n x (kb-1)
    {
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
  (void)p;
}

long long smooth_scratch_doubles(int n) {
  // Q basis + Ritz coefficient vectors + U + X (row-major) + E tiles + tile
  // partials (when they exceed the LDS work area)
  long long nu = n;
  long long s = nu * (nu + 1) + (long long)kMaxBasis * nu + nu * kMaxBasis +
                (long long)n * kMaxBasis + (long long)smooth_tile_count(n) * (4096 + 128) + 64;
  if (s < 2LL * n) s = 2LL * n;
  return (s + 31) / 32 * 32;
}

size_t smooth_lds_bytes(int n) {
  int np2 = 64;
  while (np2 < n) np2 <<= 1;
  return smooth_lds_size(np2) + (size_t)smooth_cache_cols(n) * n * sizeof(double);
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
