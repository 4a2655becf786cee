// Classical Gram-Schmidt pass of a Lanczos kernel (the thin-plate smooth
// setup): z (LDS, n rows) against the basis columns 0 .. nb - 1 of Q (device
// memory, column b at Q + b * ld).  (The dcov component kernel keeps its own
// two-load-deep form: at its 128-VGPR budget deeper batches spill.)
#pragma once
#include "kpc_common.h"

namespace kpc {

// Sums over the wave of the eight values v[0..7] by a transposed butterfly:
// the xor-32 stage exchanges four values and keeps the half selected by lane
// bit 5, the xor-16 stage two, the xor-8 stage one, then three plain stages;
// 10 cross-lane adds instead of 8 x 6.  Lane l returns the total of value
// 4 bit5(l) + 2 bit4(l) + bit3(l).
__device__ __forceinline__ double wave_sum8_transposed(const double (&v)[8]) {
  const int lane = threadIdx.x & (KPC_WAVE - 1);
  const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8;
  double h4[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double keep = b5 ? v[4 + k] : v[k];
    const double give = b5 ? v[k] : v[4 + k];
    h4[k] = keep + __shfl_xor(give, 32, KPC_WAVE);
  }
  double h2[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double keep = b4 ? h4[2 + k] : h4[k];
    const double give = b4 ? h4[k] : h4[2 + k];
    h2[k] = keep + __shfl_xor(give, 16, KPC_WAVE);
  }
  double s = (b3 ? h2[1] : h2[0]) + __shfl_xor(b3 ? h2[0] : h2[1], 8, KPC_WAVE);
  s += __shfl_xor(s, 4, KPC_WAVE);
  s += __shfl_xor(s, 2, KPC_WAVE);
  s += __shfl_xor(s, 1, KPC_WAVE);
  return s;
}

// One pass: coef[b] = Q[:, b]' z for b < nb (a wave per eight columns, lanes
// over rows, RB row blocks = 8 RB loads in flight; basis columns past nb - 1
// are clamped to it so every load is unconditional), then z -= Q coef row by
// row with eight accumulators (SB loads in flight, SB a multiple of 8).
// Returns |z|^2 after the pass.  All threads must call it; red: block-sum
// scratch.  The batch depths change only how many loads are in flight, never
// the arithmetic: callers size them to their register budget.
template <int RB, int SB>
__device__ inline double cgs_pass(const double* __restrict__ Q, long long ld, int n, int nb,
                                  double* z, double* coef, double* red) {
  const int tid = threadIdx.x;
  const int lane = tid & (KPC_WAVE - 1);
  const int wid = tid / KPC_WAVE;
  const int nw = blockDim.x / KPC_WAVE;
  for (int b0 = wid * 8; b0 < nb; b0 += nw * 8) {
    const int nc = min(8, nb - b0);
    const double* qcol[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) qcol[c] = Q + (size_t)(b0 + (c < nc ? c : nc - 1)) * ld;
    double g[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int i = lane;
    for (; i + (RB - 1) * KPC_WAVE < n; i += RB * KPC_WAVE) {
      double v[RB][8], zz[RB];
#pragma unroll
      for (int u = 0; u < RB; ++u) {
#pragma unroll
        for (int c = 0; c < 8; ++c) v[u][c] = qcol[c][i + u * KPC_WAVE];
        zz[u] = z[i + u * KPC_WAVE];
      }
#pragma unroll
      for (int u = 0; u < RB; ++u)
#pragma unroll
        for (int c = 0; c < 8; ++c) g[c] += zz[u] * v[u][c];
    }
    for (; i < n; i += KPC_WAVE) {
      double v[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) v[c] = qcol[c][i];
      const double zi = z[i];
#pragma unroll
      for (int c = 0; c < 8; ++c) g[c] += zi * v[c];
    }
    const double s = wave_sum8_transposed(g);
    const int c = ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1);
    if ((lane & 7) == 0 && c < nc) coef[b0 + c] = s;
  }
  __syncthreads();
  double part = 0.0;
  for (int i = tid; i < n; i += blockDim.x) {
    double sa[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int b = 0;
    for (; b + SB <= nb; b += SB) {
      double v[SB];
#pragma unroll
      for (int u = 0; u < SB; ++u) v[u] = Q[(size_t)(b + u) * ld + i];
#pragma unroll
      for (int u = 0; u < SB; ++u) sa[u & 7] += coef[b + u] * v[u];
    }
    for (; b + 8 <= nb; b += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = Q[(size_t)(b + u) * ld + i];
#pragma unroll
      for (int u = 0; u < 8; ++u) sa[u] += coef[b + u] * v[u];
    }
    for (; b < nb; ++b) sa[0] += coef[b] * Q[(size_t)b * ld + i];
    const double zi =
        z[i] - (((sa[0] + sa[1]) + (sa[2] + sa[3])) + ((sa[4] + sa[5]) + (sa[6] + sa[7])));
    z[i] = zi;
    part += zi * zi;
  }
  return block_sum(part, red);
}

}  // namespace kpc
