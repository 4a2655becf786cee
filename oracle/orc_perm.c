/*
 * Permutation tests for the CPU oracle (test infrastructure only).
 *
 * Permutation indices.  The reference draws per-task permutation tables on the
 * host (R's RNG stream, fastkpc/src/full_cuda_ci_one_call.cpp:464-567, or
 * seeded std::mt19937 + std::shuffle per task,
 * fastkpc/src/cuda/full_cuda_ci_method_batch.cu:1681-1717 with the task seed of
 * fastkpc/src/full_cuda_ci_one_call.cpp:296-325).  This build replaces both with
 * a counter-based generator that the GPU evaluates in place and the oracle
 * evaluates identically, so both consume the same index arrays:
 *   - task seed: SplitMix64 finaliser (Steele, Lea & Flood 2014) folded over
 *     (user seed, method, min(x,y), max(x,y), |S|, S sorted);
 *   - draw i of replicate r: Philox4x32-10 (Salmon et al., SC'11) with key =
 *     task seed and counter {i >> 1, r, 0x4b504331, 0}; the 64-bit value is
 *     words (1,0) for even i and (3,2) for odd i;
 *   - permutation: Durstenfeld shuffle of the identity, i = n-1 .. 1, swapping
 *     positions i and floor(u * (i + 1) / 2^64).
 *
 * dcc.perm (energy::dcov.test as run by the reference,
 * fastkpc/src/cuda/full_cuda_ci_method_batch.cu:284-443): double-centred
 * distance matrices A, B; observed sum T = sum A_ij B_ij, replicate sums
 * sum A_ij B_{pi(i) pi(j)}; exceedances start at include_observed and count
 * replicate >= T; statistic T / n; p = exceed / (R + include_observed), p = 1
 * when T <= 0.
 */
#include "kpc_oracle.h"
#include "orc_internal.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void orc_philox4x32_10(const unsigned int ctr[4], const unsigned int key[2], unsigned int out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    if (round > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static uint64_t splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

unsigned long long orc_task_seed(unsigned long long user_seed, int method, int x, int y,
                                 const int* S, int s_len) {
  const int a = x < y ? x : y, b = x < y ? y : x;
  uint64_t h = splitmix((uint64_t)user_seed ^ (uint64_t)(method + 1));
  h = splitmix(h ^ (uint64_t)(a + 1));
  h = splitmix(h ^ (uint64_t)(b + 1));
  h = splitmix(h ^ (uint64_t)(s_len + 1));
  for (int i = 0; i < s_len; ++i) h = splitmix(h ^ (uint64_t)(S[i] + 1));
  return h;
}

void orc_permutation(unsigned long long seed, int replicate, int n, int* perm) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int i = n - 1; i >= 1; --i) {
    const uint32_t ctr[4] = {(uint32_t)(i >> 1), (uint32_t)replicate, 0x4b504331u, 0u};
    uint32_t w[4];
    orc_philox4x32_10(ctr, key, w);
    const uint64_t u = (i & 1) ? ((uint64_t)w[3] << 32 | w[2]) : ((uint64_t)w[1] << 32 | w[0]);
    const int j = (int)(((unsigned __int128)u * (unsigned __int128)(uint64_t)(i + 1)) >> 64);
    const int t = perm[i];
    perm[i] = perm[j];
    perm[j] = t;
  }
}

/* Double-centred distance matrix A = D - rowmean - colmean + grandmean with
 * D_ij = |x_i - x_j|^index (fastkpc/src/cuda/full_cuda_ci_method_batch.cu:
 * 284-339). */
static void centred_distance(const double* x, int n, double index, double* a) {
  double* rs = (double*)malloc(sizeof(double) * n);
  double total = 0.0;
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) {
      double d = fabs(x[i] - x[j]);
      if (index != 1.0) d = pow(d, index);
      a[(size_t)i * n + j] = d;
      s += d;
    }
    rs[i] = s;
  }
  for (int i = 0; i < n; ++i) total += rs[i];
  const double nd = (double)n;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      a[(size_t)i * n + j] = a[(size_t)i * n + j] - rs[i] / nd - rs[j] / nd + total / (nd * nd);
  free(rs);
}

/* Shared tail of the permutation tests: from the observed value and the
 * replicate values, out = {p, statistic, mean(rep), var(rep), margin}. */
static void perm_finish(double observed, double stat_scale, const double* rep, int R,
                        int include_observed, int p_one_if_nonpositive, double out[5]) {
  int exceed = include_observed ? 1 : 0;
  double margin = 1e300;
  double mean = 0.0, m2 = 0.0;
  for (int r = 0; r < R; ++r) {
    if (rep[r] >= observed) ++exceed;
    const double g = fabs(rep[r] - observed) / (fabs(observed) > 0.0 ? fabs(observed) : 1.0);
    if (g < margin) margin = g;
    const double v = rep[r] * stat_scale;
    const double d = v - mean;
    mean += d / (double)(r + 1);
    m2 += d * (v - mean);
  }
  const int denom = R + (include_observed ? 1 : 0);
  double p = denom > 0 ? (double)exceed / (double)denom : 1.0;
  if (p_one_if_nonpositive && !(observed > 0.0)) p = 1.0;
  out[0] = p;
  out[1] = observed * stat_scale;
  out[2] = mean;
  out[3] = R > 1 ? m2 / (double)(R - 1) : 0.0;
  out[4] = margin;
}

/* dcc.perm with replicate r permuting y by table[r] (R x n, 0-based; the
 * permuted sample is y[table[r][i]]); table NULL: the counter-based indices
 * of `seed` (orc_permutation). */
int orc_dcc_perm_table(const double* x, const double* y, int n, double index, int R,
                       int include_observed, unsigned long long seed, const int* table,
                       double out[5]) {
  if (n < 2 || R < 0) return -1;
  double* a = (double*)malloc(sizeof(double) * (size_t)n * n);
  double* b = (double*)malloc(sizeof(double) * (size_t)n * n);
  double* rep = (double*)malloc(sizeof(double) * (R > 0 ? R : 1));
  int* own = (int*)malloc(sizeof(int) * n);
  centred_distance(x, n, index, a);
  centred_distance(y, n, index, b);
  double observed = 0.0;
  for (size_t c = 0; c < (size_t)n * n; ++c) observed += a[c] * b[c];
  for (int r = 0; r < R; ++r) {
    const int* perm = own;
    if (table) perm = table + (size_t)r * n;
    else orc_permutation(seed, r, n, own);
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
      const double* br = b + (size_t)perm[i] * n;
      for (int j = 0; j < n; ++j) s += a[(size_t)i * n + j] * br[perm[j]];
    }
    rep[r] = s;
  }
  perm_finish(observed, 1.0 / (double)n, rep, R, include_observed, 1, out);
  free(a); free(b); free(rep); free(own);
  return 0;
}

int orc_dcc_perm(const double* x, const double* y, int n, double index, int R,
                 int include_observed, unsigned long long seed, double out[5]) {
  return orc_dcc_perm_table(x, y, n, index, R, include_observed, seed, NULL, out);
}

/* hsic.perm on precomputed centred factors (kpcalg/R/hsicperm.R:77-88; the
 * rows of the second factor are permuted by table[r], or by the counter-based
 * indices of `seed` when table is NULL). */
void orc_hsic_perm_table(const double* gx, int rx, const double* gy, int ry, int n, int R,
                         int include_observed, unsigned long long seed, const int* table,
                         double out[5]) {
  if (n < 1) return;
  double* rep = (double*)malloc(sizeof(double) * (R > 0 ? R : 1));
  int* own = (int*)malloc(sizeof(int) * n);
  int* ident = (int*)malloc(sizeof(int) * n);
  for (int i = 0; i < n; ++i) ident[i] = i;
  const double n2 = (double)n * (double)n;
  double observed = orc_hsic_cross(gx, rx, gy, ry, n, ident) / n2;
  for (int r = 0; r < R; ++r) {
    const int* perm = own;
    if (table) perm = table + (size_t)r * n;
    else orc_permutation(seed, r, n, own);
    rep[r] = orc_hsic_cross(gx, rx, gy, ry, n, perm) / n2;
  }
  perm_finish(observed, 1.0, rep, R, include_observed, 0, out);
  free(rep); free(own); free(ident);
}

void orc_hsic_perm_from_factors(const double* gx, int rx, const double* gy, int ry, int n,
                                int R, int include_observed, unsigned long long seed,
                                double out[5]) {
  orc_hsic_perm_table(gx, rx, gy, ry, n, R, include_observed, seed, NULL, out);
}
