// Permutation tests (dcc.perm, hsic.perm) for gfx950.
//
// Semantics (fastkpc/src/cuda/full_cuda_ci_method_batch.cu:341-443, 791-908;
// kpcalg/R/hsicperm.R:77-88; energy::dcov.test): replicate r of test t
// evaluates the statistic with the second variable's samples permuted by the
// r-th index array of the test's seed; p = (#{T_r >= T} + include_observed) /
// (R + include_observed).  The index arrays are the Durstenfeld shuffles of
// 0..n-1 driven by Philox4x32-10 (key = task seed, counter {i >> 1, r,
// 0x4b504331, 0}), the arrays the oracle draws.
//
// Pipeline per batch of tests (host: kpc_host.cpp, Engine::perm_pairs):
//   observed   the statistic of every test, computed by the replicate kernel
//              with the identity permutation (bit-identical to a replicate
//              whose permutation is the identity);
//   per chunk of replicates (bounded scratch):
//     perm_generate_kernel  one shuffle per lane in LDS, copied out coalesced;
//     replicate kernel      one value per replicate into `vals`;
//     perm_tally_kernel     exceedances, running mean / M2 (Welford, in
//                           replicate order) and the smallest relative
//                           distance to the observed value, per test;
//   perm_finalize_kernel   p, statistic, mean, variance, margin.
//
// hsic.perm replicate (hsic_perm_kernel): |Gx' P_r Gy|_F^2 / n^2 with the
// column-centred inchol factors stored row-major and zero-padded to 16 KT
// columns.  One wave per replicate forms the KT x KT grid of 16 x 16 output
// tiles with v_mfma_f64_16x16x4f64 over the samples: lane (g, c) = (lane >> 4,
// lane & 15) feeds A[i][16 ta + c] and B[perm(i)][16 tb + c] for sample
// i = i0 + g, so each 16-lane group reads one 128-byte row segment.
//
// dcc.perm replicate (dcc_perm_kernel): with A the double-centred distance
// matrix of x (zero row and column sums), sum_ij A_ij B'_ij equals
// sum_ij A_ij |y'_i - y'_j|^index for the permuted y' -- the centring of B
// drops out.  A workgroup owns one test and a block of RB replicates; a wave
// takes 128 x JT tiles of the upper triangle (lane = rows i and i + 64), keeps
// the weighted A entries (2 off the diagonal, 1 on it, 0 below) in registers,
// stages the JT permuted y_j of every replicate in wave-private LDS and
// accumulates a_ij |y'_i - y'_j| with one subtract and one FMA per entry and
// replicate.  The permutation indices of the next tile are loaded while the
// current tile computes.  (JT, RB) is a launch-time tile shape.
#include "kpc_common.h"
#include "kpc_device.h"

namespace kpc {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kGenLdsBytes = 80 * 1024;

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

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

// One wave per unit: lane l < nr shuffles replicate r0 + l in its own LDS
// row, then the wave copies the unit's rows to perms[slot .. slot + nr).
__global__ __launch_bounds__(KPC_WAVE) void perm_generate_kernel(
    const PermUnit* __restrict__ units, int nunits, const unsigned long long* __restrict__ seeds,
    int n, unsigned short* __restrict__ perms) {
  extern __shared__ unsigned short pl[];
  const int lane = threadIdx.x;
  for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
    const PermUnit U = units[u];
    if (lane < U.nr) make_permutation<unsigned short>(seeds[U.test], U.r0 + lane, n, pl + (size_t)lane * n);
    __syncthreads();
    unsigned short* dst = perms + (size_t)U.slot * n;
    const int total = U.nr * n;
    for (int e = lane; e < total; e += KPC_WAVE) dst[e] = pl[e];
    __syncthreads();
  }
}

// ------------------------------------------------------------- hsic.perm
// sum over the KT x KT output tiles of (sum_i A[i][a] B[perm(i)][b])^2 for one
// wave; perm == nullptr is the identity.
template <int KT>
__device__ __forceinline__ double hsic_cross_sumsq(const double* __restrict__ A,
                                                   const double* __restrict__ B, int n,
                                                   const unsigned short* __restrict__ perm) {
  constexpr int kp = 16 * KT;
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int g = lane >> 4, c = lane & 15;
  v4d acc[KT][KT];
#pragma unroll
  for (int a = 0; a < KT; ++a)
#pragma unroll
    for (int b = 0; b < KT; ++b) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
  for (int i0 = 0; i0 < n; i0 += KPC_WAVE) {
    const int il = i0 + lane;
    const int pv = il < n ? (perm ? (int)perm[il] : il) : 0;
    const int steps = min(16, (n - i0 + 3) >> 2);
#pragma unroll 4
    for (int s = 0; s < steps; ++s) {
      const int ii = 4 * s + g;
      const int i = i0 + ii;
      const int pi = __shfl(pv, ii, KPC_WAVE);
      const bool ok = i < n;
      const int ia = ok ? i : 0;
      double av[KT], bv[KT];
#pragma unroll
      for (int t = 0; t < KT; ++t) {
        const double x = A[(size_t)ia * kp + 16 * t + c];
        const double y = B[(size_t)pi * kp + 16 * t + c];
        av[t] = ok ? x : 0.0;
        bv[t] = ok ? y : 0.0;
      }
#pragma unroll
      for (int a = 0; a < KT; ++a)
#pragma unroll
        for (int b = 0; b < KT; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
  }
  double ss = 0.0;
#pragma unroll
  for (int a = 0; a < KT; ++a)
#pragma unroll
    for (int b = 0; b < KT; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) ss += acc[a][b][v] * acc[a][b][v];
  return wave_sum(ss);
}

template <int KT>
__global__ __launch_bounds__(256) void hsic_perm_kernel(
    const PermUnit* __restrict__ units, int nunits, const int* __restrict__ pa,
    const int* __restrict__ pb, int n, const double* __restrict__ rows,
    const unsigned short* __restrict__ perms, double* __restrict__ vals) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int wid = threadIdx.x / KPC_WAVE;
  const int nw = blockDim.x / KPC_WAVE;
  const double nd = (double)n;
  const size_t cstride = (size_t)n * 16 * KT;
  for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
    const PermUnit U = units[u];
    const double* A = rows + (size_t)pa[U.test] * cstride;
    const double* B = rows + (size_t)pb[U.test] * cstride;
    for (int r = wid; r < U.nr; r += nw) {
      const unsigned short* pr = perms ? perms + (size_t)(U.slot + r) * n : nullptr;
      const double v = hsic_cross_sumsq<KT>(A, B, n, pr);
      if (lane == 0) vals[U.slot + r] = v / (nd * nd);
    }
  }
}

// -------------------------------------------------------------- dcc.perm
template <bool kUnitIndex>
__device__ __forceinline__ double dist_pow(double d, double index) {
  return kUnitIndex ? fabs(d) : pow(fabs(d), index);
}

// Tiles of the upper triangle: row blocks of 128 (two rows per lane), column
// blocks of JT with JT cb + JT - 1 >= 128 rb.
template <int JT>
__device__ __forceinline__ int dcc_tile_count(int n) {
  constexpr int kc = 128 / JT;
  const int nrb = (n + 127) / 128, ncb = (n + JT - 1) / JT;
  int t = 0;
  for (int rb = 0; rb < nrb; ++rb) t += max(0, ncb - kc * rb);
  return t;
}

// Replicate values sum_ij a_ij |y'_i - y'_j|^index for the RB replicates of
// each unit.  LDS: XR = {x, row mean} per sample (both doubled for index 1,
// so the doubled off-diagonal weight is exact: (|2x_i - 2x_j| + (2abar -
// 2r_i)) - 2r_j), the second variable y, and per wave RB x JT staged y'_j.
// Replicate r of a unit only ever meets replicate r's own values, in the
// same order for every RB, so RB = 1 (the observed pass) reproduces a
// replicate with the identity permutation bit for bit.
template <bool kUnitIndex, int JT, int RB>
__global__ __launch_bounds__(256) void dcc_perm_kernel(
    const PermUnit* __restrict__ units, int nunits, const int* __restrict__ pa,
    const int* __restrict__ pb, int n, int stride, double index, const double* __restrict__ vecs,
    const double* __restrict__ grand, const unsigned short* __restrict__ perms,
    double* __restrict__ vals) {
  constexpr int kc = 128 / JT;
  constexpr int kStage = (RB * JT + KPC_WAVE - 1) / KPC_WAVE;  // staged entries per lane
  constexpr unsigned short kNone = 0xFFFF;                      // entry outside the data: 0
  extern __shared__ double dl[];
  __shared__ double red[4 * RB];
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int wid = threadIdx.x / KPC_WAVE;
  const int nw = blockDim.x / KPC_WAVE;
  double2* XR = (double2*)dl;
  double* Y = dl + 2 * n;
  double* stage = dl + ((3 * n + 1) & ~1) + (size_t)wid * RB * JT;
  const double2* stage2 = (const double2*)stage;
  const double sc = kUnitIndex ? 2.0 : 1.0;
  const int ncb = (n + JT - 1) / JT;
  const int ntiles = dcc_tile_count<JT>(n);
  for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
    const PermUnit U = units[u];
    const double* VA = vecs + (size_t)pa[U.test] * stride;
    const double* VB = vecs + (size_t)pb[U.test] * stride;
    const double abar = grand[pa[U.test]];
    const double abar_s = sc * abar;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      XR[i] = make_double2(sc * VA[i], sc * VA[n + i]);
      Y[i] = VB[i];
    }
    __syncthreads();
    auto decode = [&](int t, int& rb, int& cb) {
      rb = 0;
      int rem = t;
      while (rem >= ncb - kc * rb) {
        rem -= ncb - kc * rb;
        ++rb;
      }
      cb = kc * rb + rem;
    };
    // permutation indices of one tile: rows i0, i0 + 64 and the staged columns
    unsigned short ix0[RB], ix1[RB], ixs[kStage];
    auto fetch = [&](int t) {
      int rb, cb;
      decode(t, rb, cb);
      const int i0 = 128 * rb + lane, i1 = i0 + 64, j0 = JT * cb;
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const unsigned short* pr = perms ? perms + (size_t)(U.slot + r) * n : nullptr;
        const bool live = r < U.nr;
        ix0[r] = (live && i0 < n) ? (pr ? pr[i0] : (unsigned short)i0) : kNone;
        ix1[r] = (live && i1 < n) ? (pr ? pr[i1] : (unsigned short)i1) : kNone;
      }
#pragma unroll
      for (int k = 0; k < kStage; ++k) {
        const int e = lane + KPC_WAVE * k;
        const int r = e / JT, j = j0 + (e - (e / JT) * JT);
        unsigned short v = kNone;
        if (e < RB * JT && r < U.nr && j < n)
          v = perms ? perms[(size_t)(U.slot + r) * n + j] : (unsigned short)j;
        ixs[k] = v;
      }
    };
    double acc0[RB], acc1[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) acc0[r] = acc1[r] = 0.0;
    if (wid < ntiles) fetch(wid);
    for (int t = wid; t < ntiles; t += nw) {
      int rb, cb;
      decode(t, rb, cb);
      const int i0 = 128 * rb + lane, i1 = i0 + 64, j0 = JT * cb;
      // stage the permuted values of this tile, then prefetch the next tile's indices
      double y0[RB], y1[RB];
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        y0[r] = ix0[r] == kNone ? 0.0 : Y[ix0[r]];
        y1[r] = ix1[r] == kNone ? 0.0 : Y[ix1[r]];
      }
#pragma unroll
      for (int k = 0; k < kStage; ++k) {
        const int e = lane + KPC_WAVE * k;
        if (e < RB * JT) stage[e] = ixs[k] == kNone ? 0.0 : Y[ixs[k]];
      }
      if (t + nw < ntiles) fetch(t + nw);
      // weighted centred distances of rows i0, i1 against the tile's columns
      const bool v0 = i0 < n, v1 = i1 < n;
      const double2 x0 = XR[v0 ? i0 : 0];
      const double2 x1 = XR[v1 ? i1 : 0];
      const double c0 = abar_s - x0.y, c1 = abar_s - x1.y;
      double a0[JT], a1[JT];
      const bool interior = 128 * rb + 127 < n && j0 + JT - 1 < n && j0 > 128 * rb + 127;
      if (interior) {
#pragma unroll
        for (int jj = 0; jj < JT; ++jj) {
          const double2 xj = XR[j0 + jj];
          if (kUnitIndex) {
            a0[jj] = (fabs(x0.x - xj.x) + c0) - xj.y;
            a1[jj] = (fabs(x1.x - xj.x) + c1) - xj.y;
          } else {
            a0[jj] = 2.0 * ((dist_pow<false>(x0.x - xj.x, index) + c0) - xj.y);
            a1[jj] = 2.0 * ((dist_pow<false>(x1.x - xj.x, index) + c1) - xj.y);
          }
        }
      } else {
#pragma unroll
        for (int jj = 0; jj < JT; ++jj) {
          const int j = j0 + jj;
          const bool vj = j < n;
          const double2 xj = XR[vj ? j : 0];
          double e0 = (dist_pow<kUnitIndex>(x0.x - xj.x, index) + c0) - xj.y;
          double e1 = (dist_pow<kUnitIndex>(x1.x - xj.x, index) + c1) - xj.y;
          if (!kUnitIndex) {
            e0 *= 2.0;
            e1 *= 2.0;
          }
          a0[jj] = (v0 && vj && j >= i0) ? (j > i0 ? e0 : 0.5 * e0) : 0.0;
          a1[jj] = (v1 && vj && j >= i1) ? (j > i1 ? e1 : 0.5 * e1) : 0.0;
        }
      }
      wave_lds_sync();
#pragma unroll
      for (int jp = 0; jp < JT / 2; ++jp) {
#pragma unroll
        for (int r = 0; r < RB; ++r) {
          const double2 yj = stage2[r * (JT / 2) + jp];
          acc0[r] = fma(a0[2 * jp], dist_pow<kUnitIndex>(y0[r] - yj.x, index), acc0[r]);
          acc1[r] = fma(a1[2 * jp], dist_pow<kUnitIndex>(y1[r] - yj.x, index), acc1[r]);
          acc0[r] = fma(a0[2 * jp + 1], dist_pow<kUnitIndex>(y0[r] - yj.y, index), acc0[r]);
          acc1[r] = fma(a1[2 * jp + 1], dist_pow<kUnitIndex>(y1[r] - yj.y, index), acc1[r]);
        }
      }
      wave_lds_sync();
    }
    // weight 2 off the diagonal covers (i, j) and (j, i): vals hold the sum
    // over the full matrix
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const double s = wave_sum(acc0[r] + acc1[r]);
      if (lane == 0) red[wid * RB + r] = s;
    }
    __syncthreads();
    if ((int)threadIdx.x < U.nr) {
      double tot = 0.0;
      for (int w = 0; w < nw; ++w) tot += red[w * RB + threadIdx.x];
      vals[U.slot + threadIdx.x] = tot;
    }
  }
}

// ------------------------------------------------------------ tally / p
// One wave per segment (a test's replicates r0 .. r0 + nr, in order); the
// state {exceedances, mean, M2, margin} of the test carries across chunks.
__global__ __launch_bounds__(256) void perm_tally_kernel(const PermUnit* __restrict__ segs,
                                                         int nsegs, const double* __restrict__ vals,
                                                         const double* __restrict__ obs,
                                                         double scale, double* __restrict__ state) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int w = blockIdx.x * (blockDim.x / KPC_WAVE) + threadIdx.x / KPC_WAVE;
  if (w >= nsegs) return;
  const PermUnit U = segs[w];
  double* st = state + 4 * (size_t)U.test;
  double exceed = st[0], mean = st[1], m2 = st[2], margin = st[3];
  const double observed = obs[U.test];
  const double denom = fabs(observed) > 0.0 ? fabs(observed) : 1.0;
  for (int b = 0; b < U.nr; b += KPC_WAVE) {
    const double mine = b + lane < U.nr ? vals[U.slot + b + lane] : 0.0;
    const int cnt = min(KPC_WAVE, U.nr - b);
    for (int e = 0; e < cnt; ++e) {
      const double v = __shfl(mine, e, KPC_WAVE);
      if (lane == 0) {
        if (v >= observed) exceed += 1.0;
        margin = fmin(margin, fabs(v - observed) / denom);
        const double s = v * scale;
        const double d = s - mean;
        mean += d / (double)(U.r0 + b + e + 1);
        m2 += d * (s - mean);
      }
    }
  }
  if (lane == 0) {
    st[0] = exceed;
    st[1] = mean;
    st[2] = m2;
    st[3] = margin;
  }
}

__global__ void perm_finalize_kernel(int count, const double* __restrict__ obs,
                                     const double* __restrict__ state, double scale,
                                     int replicates, int include_observed, int one_if_nonpositive,
                                     double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const double* st = state + 4 * (size_t)t;
  const double observed = obs[t];
  const int denom = replicates + (include_observed ? 1 : 0);
  double p = denom > 0 ? st[0] / (double)denom : 1.0;
  if (one_if_nonpositive && !(observed > 0.0)) p = 1.0;
  double* o = out + (size_t)t * kPairOut;
  o[0] = p;
  o[1] = observed * scale;
  o[2] = st[1];
  o[3] = replicates > 1 ? st[2] / (double)(replicates - 1) : 0.0;
  o[4] = st[3];
}

// ---------------------------------------------------------------- launchers
int perm_generate_width(int n) {
  int g = kGenLdsBytes / (2 * n);
  return g < 1 ? 1 : (g > KPC_WAVE ? KPC_WAVE : g);
}

hipError_t launch_perm_generate(const PermUnit* units, int nunits,
                                const unsigned long long* seeds, int n, int width,
                                unsigned short* perms, int grid, hipStream_t stream) {
  if (nunits <= 0) return hipSuccess;
  if (width < 1 || width > KPC_WAVE) return hipErrorInvalidValue;
  const size_t lds = (size_t)width * n * sizeof(unsigned short);
  hipError_t err = hipFuncSetAttribute((const void*)perm_generate_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(perm_generate_kernel, dim3(grid), dim3(KPC_WAVE), lds, stream, units, nunits,
                     seeds, n, perms);
  return hipGetLastError();
}

hipError_t launch_hsic_perm_values(const PermUnit* units, int nunits, const int* pa,
                                   const int* pb, int n, int kt, const double* rows,
                                   const unsigned short* perms, double* vals, int grid,
                                   hipStream_t stream) {
  if (nunits <= 0) return hipSuccess;
  switch (kt) {
    case 1:
      hipLaunchKernelGGL(hsic_perm_kernel<1>, dim3(grid), dim3(256), 0, stream, units, nunits, pa,
                         pb, n, rows, perms, vals);
      break;
    case 2:
      hipLaunchKernelGGL(hsic_perm_kernel<2>, dim3(grid), dim3(256), 0, stream, units, nunits, pa,
                         pb, n, rows, perms, vals);
      break;
    case 3:
      hipLaunchKernelGGL(hsic_perm_kernel<3>, dim3(grid), dim3(256), 0, stream, units, nunits, pa,
                         pb, n, rows, perms, vals);
      break;
    case 4:
      hipLaunchKernelGGL(hsic_perm_kernel<4>, dim3(grid), dim3(256), 0, stream, units, nunits, pa,
                         pb, n, rows, perms, vals);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

size_t dcc_perm_lds_bytes(int n, int rb, int jt) {
  return ((size_t)((3 * n + 1) & ~1) + (size_t)4 * rb * jt) * sizeof(double);
}

template <bool kUnit, int JT, int RB>
static hipError_t dcc_launch(const PermUnit* units, int nunits, const int* pa, const int* pb,
                             int n, int stride, double index, const double* vecs,
                             const double* grand, const unsigned short* perms, double* vals,
                             int grid, hipStream_t stream) {
  const size_t lds = dcc_perm_lds_bytes(n, RB, JT);
  hipError_t err = hipFuncSetAttribute((const void*)dcc_perm_kernel<kUnit, JT, RB>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL((dcc_perm_kernel<kUnit, JT, RB>), dim3(grid), dim3(256), lds, stream, units,
                     nunits, pa, pb, n, stride, index, vecs, grand, perms, vals);
  return hipGetLastError();
}

template <int JT, int RB>
static hipError_t dcc_dispatch(bool unit, const PermUnit* units, int nunits, const int* pa,
                               const int* pb, int n, int stride, double index,
                               const double* vecs, const double* grand,
                               const unsigned short* perms, double* vals, int grid,
                               hipStream_t stream) {
  return unit ? dcc_launch<true, JT, RB>(units, nunits, pa, pb, n, stride, index, vecs, grand,
                                         perms, vals, grid, stream)
              : dcc_launch<false, JT, RB>(units, nunits, pa, pb, n, stride, index, vecs, grand,
                                          perms, vals, grid, stream);
}

// Tile shapes (columns per tile JT, replicates per unit RB); RB = 1 is the
// observed pass of any JT.
bool dcc_perm_shape_supported(int jt, int rb) {
  return (rb == 1 && (jt == 8 || jt == 16 || jt == 32)) ||
         (jt == 16 && (rb == 8 || rb == 12)) || (jt == 8 && rb == 16) || (jt == 32 && rb == 8);
}

hipError_t launch_dcc_perm_values(const PermUnit* units, int nunits, const int* pa, const int* pb,
                                  int n, int stride, double index, const double* vecs,
                                  const double* grand, const unsigned short* perms, double* vals,
                                  int jt, int rb, int grid, hipStream_t stream) {
  if (nunits <= 0) return hipSuccess;
  const bool unit = index == 1.0;
#define KPC_DCC_SHAPE(J, B)                                                                    \
  if (jt == J && rb == B)                                                                      \
    return dcc_dispatch<J, B>(unit, units, nunits, pa, pb, n, stride, index, vecs, grand, perms, \
                              vals, grid, stream);
  KPC_DCC_SHAPE(8, 1)
  KPC_DCC_SHAPE(16, 1)
  KPC_DCC_SHAPE(32, 1)
  KPC_DCC_SHAPE(16, 8)
  KPC_DCC_SHAPE(16, 12)
  KPC_DCC_SHAPE(8, 16)
  KPC_DCC_SHAPE(32, 8)
#undef KPC_DCC_SHAPE
  return hipErrorInvalidValue;
}

hipError_t launch_perm_tally(const PermUnit* segs, int nsegs, const double* vals,
                             const double* obs, double scale, double* state, hipStream_t stream) {
  if (nsegs <= 0) return hipSuccess;
  const int waves = 4;
  hipLaunchKernelGGL(perm_tally_kernel, dim3((nsegs + waves - 1) / waves), dim3(waves * KPC_WAVE),
                     0, stream, segs, nsegs, vals, obs, scale, state);
  return hipGetLastError();
}

hipError_t launch_perm_finalize(int count, const double* obs, const double* state, double scale,
                                int replicates, int include_observed, int one_if_nonpositive,
                                double* out, hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(perm_finalize_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, count,
                     obs, state, scale, replicates, include_observed, one_if_nonpositive, out);
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
