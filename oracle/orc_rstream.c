/*
 * R's global random stream and the permutation tables kpcalg's permutation
 * tests draw from it (test infrastructure only).
 *
 * Restates, in plain C: R's Mersenne-Twister kind as seeded by set.seed()
 * (RNG.c: 50 scrambling rounds of seed = 69069 seed + 1, one more per state
 * word, position reset to 624), unif_rand() (tempered word times 2^-32, kept
 * in (0, 1)), R_unif_index() under the "Rejection" sample.kind of R >= 3.6
 * (rbits: 16 bits per uniform, masked to ceil(log2 n) bits, redrawn while
 * >= n) and sample(n); energy's permute() used by dcov.test (n - 1 swaps
 * with j = floor(m u)); and the order fastkpc draws them in
 * (fastkpc/src/full_cuda_ci_one_call.cpp:464-567: dcc.perm keeps one index
 * vector per test across its replicates, hsic.perm draws sample(n) per
 * replicate).  A "Rounding" sample.kind (floor(n u), R < 3.6) is provided
 * only to check energy's documented identity permute(J) == rev(sample(J)).
 */
#include "kpc_oracle.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

#define ORC_MT_N 624
#define ORC_MT_M 397

void orc_rng_set_seed(orc_rng* r, unsigned int seed) {
  uint32_t s = seed;
  for (int j = 0; j < 50; ++j) s = 69069u * s + 1u;
  for (int j = 0; j < ORC_MT_N + 1; ++j) {
    s = 69069u * s + 1u;
    if (j > 0) r->mt[j - 1] = s;
  }
  r->mti = ORC_MT_N;
}

unsigned int orc_rng_word(orc_rng* r) {
  static const uint32_t mag01[2] = {0x0u, 0x9908b0dfu};
  uint32_t y;
  if (r->mti >= ORC_MT_N) {
    int kk;
    for (kk = 0; kk < ORC_MT_N - ORC_MT_M; kk++) {
      y = (r->mt[kk] & 0x80000000u) | (r->mt[kk + 1] & 0x7fffffffu);
      r->mt[kk] = r->mt[kk + ORC_MT_M] ^ (y >> 1) ^ mag01[y & 0x1u];
    }
    for (; kk < ORC_MT_N - 1; kk++) {
      y = (r->mt[kk] & 0x80000000u) | (r->mt[kk + 1] & 0x7fffffffu);
      r->mt[kk] = r->mt[kk + (ORC_MT_M - ORC_MT_N)] ^ (y >> 1) ^ mag01[y & 0x1u];
    }
    y = (r->mt[ORC_MT_N - 1] & 0x80000000u) | (r->mt[0] & 0x7fffffffu);
    r->mt[ORC_MT_N - 1] = r->mt[ORC_MT_M - 1] ^ (y >> 1) ^ mag01[y & 0x1u];
    r->mti = 0;
  }
  y = r->mt[r->mti++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= (y >> 18);
  return y;
}

double orc_unif_rand(orc_rng* r) {
  const double i2_32m1 = 2.328306437080797e-10;
  double x = (double)orc_rng_word(r) * 2.3283064365386963e-10;
  if (x <= 0.0) return 0.5 * i2_32m1;
  if ((1.0 - x) <= 0.0) return 1.0 - 0.5 * i2_32m1;
  return x;
}

static double rbits(orc_rng* r, int bits) {
  int64_t v = 0;
  for (int n = 0; n <= bits; n += 16) {
    int v1 = (int)floor(orc_unif_rand(r) * 65536);
    v = 65536 * v + v1;
  }
  return (double)(v & (((int64_t)1 << bits) - 1));
}

double orc_unif_index(orc_rng* r, double dn, int rounding) {
  if (rounding) return floor(dn * orc_unif_rand(r));
  if (dn <= 0) return 0.0;
  int bits = (int)ceil(log2(dn));
  double dv;
  do {
    dv = rbits(r, bits);
  } while (dn <= dv);
  return dv;
}

void orc_r_sample(orc_rng* r, int n, int rounding, int* pool, int* out) {
  for (int i = 0; i < n; ++i) pool[i] = i;
  int m = n;
  for (int i = 0; i < n; ++i) {
    int j = (int)orc_unif_index(r, (double)m, rounding);
    out[i] = pool[j];
    pool[j] = pool[--m];
  }
}

void orc_energy_permute(orc_rng* r, int* J, int n) {
  int m = n;
  for (int i = 0; i < n - 1; i++) {
    int j = (int)(m * orc_unif_rand(r));
    int j0 = J[j];
    J[j] = J[--m];
    J[m] = j0;
  }
}

void orc_r_perm_table(orc_rng* r, int method, int n, int R, int* pool, int* table) {
  if (method == ORC_DCC_PERM) {
    int* J = pool;
    for (int i = 0; i < n; ++i) J[i] = i;
    for (int rep = 0; rep < R; ++rep) {
      orc_energy_permute(r, J, n);
      memcpy(table + (size_t)rep * n, J, sizeof(int) * n);
    }
  } else {
    for (int rep = 0; rep < R; ++rep) orc_r_sample(r, n, 0, pool, table + (size_t)rep * n);
  }
}
