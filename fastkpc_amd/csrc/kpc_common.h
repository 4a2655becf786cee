// Shared device helpers for the gfx950 kernels of the kernel-PC CI path.
// Wavefront width is 64 on CDNA4; all reductions below assume it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define KPC_WAVE 64

namespace kpc {

constexpr int kMaxN = 2048;      // samples per column
constexpr int kMaxNumCol = 64;   // dcov/hsic truncation rank
constexpr int kMaxQ = 576;       // coefficient dimension (1 + 9 |S| with |S| <= 62)
constexpr int kMaxPen = 64;      // penalties per design (|S| for additive)

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, KPC_WAVE);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, KPC_WAVE));
  return v;
}

// Sum over the whole workgroup; every thread gets the result.  `red` must hold
// at least blockDim.x / 64 doubles and is reused after a barrier.
__device__ __forceinline__ double block_sum(double v, double* red) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int wid = threadIdx.x / KPC_WAVE;
  const int nw = (blockDim.x + KPC_WAVE - 1) / KPC_WAVE;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < nw; ++w) t += red[w];
  return t;
}

// Two workgroup sums at once (one pair of barriers); `red` must hold
// 2 * (blockDim.x / 64) doubles.
__device__ __forceinline__ void block_sum2(double& a, double& b, double* red) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int wid = threadIdx.x / KPC_WAVE;
  const int nw = (blockDim.x + KPC_WAVE - 1) / KPC_WAVE;
  a = wave_sum(a);
  b = wave_sum(b);
  __syncthreads();
  if (lane == 0) {
    red[2 * wid] = a;
    red[2 * wid + 1] = b;
  }
  __syncthreads();
  double ta = 0.0, tb = 0.0;
  for (int w = 0; w < nw; ++w) {
    ta += red[2 * w];
    tb += red[2 * w + 1];
  }
  a = ta;
  b = tb;
}

// Three workgroup sums at once (one pair of barriers); `red` must hold
// 3 * (blockDim.x / 64) doubles.
__device__ __forceinline__ void block_sum3(double& a, double& b, double& c, double* red) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int wid = threadIdx.x / KPC_WAVE;
  const int nw = (blockDim.x + KPC_WAVE - 1) / KPC_WAVE;
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  __syncthreads();
  if (lane == 0) {
    red[3 * wid] = a;
    red[3 * wid + 1] = b;
    red[3 * wid + 2] = c;
  }
  __syncthreads();
  double ta = 0.0, tb = 0.0, tc = 0.0;
  for (int w = 0; w < nw; ++w) {
    ta += red[3 * w];
    tb += red[3 * w + 1];
    tc += red[3 * w + 2];
  }
  a = ta;
  b = tb;
  c = tc;
}

__device__ __forceinline__ double block_max(double v, double* red) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int wid = threadIdx.x / KPC_WAVE;
  const int nw = (blockDim.x + KPC_WAVE - 1) / KPC_WAVE;
  v = wave_max(v);
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double t = red[0];
  for (int w = 1; w < nw; ++w) t = fmax(t, red[w]);
  return t;
}

// Error-free transformation (Knuth two-sum) for compensated sums.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Double-double workgroup sum: returns hi + lo as a pair; used where the
// reference accumulates in 80-bit long double (R's mean/colSums).
struct dd { double hi, lo; };

__device__ __forceinline__ dd dd_add(dd a, dd b) {
  double s, e;
  two_sum(a.hi, b.hi, s, e);
  e += a.lo + b.lo;
  double hi = s + e;
  double lo = e - (hi - s);
  return {hi, lo};
}

__device__ __forceinline__ dd wave_sum_dd(dd v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    dd o{__shfl_xor(v.hi, off, KPC_WAVE), __shfl_xor(v.lo, off, KPC_WAVE)};
    v = dd_add(v, o);
  }
  return v;
}

// `red` must hold 2 * (blockDim.x / 64) doubles.
__device__ __forceinline__ dd block_sum_dd(dd v, double* red) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const int wid = threadIdx.x / KPC_WAVE;
  const int nw = (blockDim.x + KPC_WAVE - 1) / KPC_WAVE;
  v = wave_sum_dd(v);
  __syncthreads();
  if (lane == 0) { red[2 * wid] = v.hi; red[2 * wid + 1] = v.lo; }
  __syncthreads();
  dd t{0.0, 0.0};
  for (int w = 0; w < nw; ++w) t = dd_add(t, dd{red[2 * w], red[2 * w + 1]});
  return t;
}

// dd / integer
__device__ __forceinline__ dd dd_div_int(dd a, double n) {
  double q1 = a.hi / n;
  // remainder a - q1*n computed exactly with fma
  double p = q1 * n;
  double pe = fma(q1, n, -p);
  double s, e;
  two_sum(a.hi, -p, s, e);
  e += a.lo - pe;
  double q2 = (s + e) / n;
  double hi = q1 + q2;
  double lo = q2 - (hi - q1);
  return {hi, lo};
}

// R's mean(): long-double sum / n, then a correction pass.  Reproduced with
// double-double arithmetic (at least as precise as x87 extended precision).
__device__ __forceinline__ double block_r_mean(const double* x, int n, double* red) {
  dd s{0.0, 0.0};
  for (int i = threadIdx.x; i < n; i += blockDim.x) s = dd_add(s, dd{x[i], 0.0});
  s = block_sum_dd(s, red);
  dd m = dd_div_int(s, (double)n);
  dd c{0.0, 0.0};
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    // (long double)x - mean, in dd
    dd d = dd_add(dd{x[i], 0.0}, dd{-m.hi, -m.lo});
    c = dd_add(c, d);
  }
  __syncthreads();
  c = block_sum_dd(c, red);
  dd corr = dd_div_int(c, (double)n);
  dd r = dd_add(m, corr);
  return r.hi + r.lo;
}

}  // namespace kpc
