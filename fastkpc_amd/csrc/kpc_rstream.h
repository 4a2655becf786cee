// R's global random stream for the permutation tests: Mersenne-Twister
// (R's default RNG kind) seeded as set.seed(seed) does, unif_rand(), and the
// index draws of the two permutation tests kpcalg calls:
//   * energy dcov.test (ci dcc.perm): one index vector per test, mutated by
//     every replicate (energy's permute(): n - 1 swaps with j = floor(m u));
//   * kpcalg hsic.perm: sample(n) per replicate (R >= 3.6 "Rejection"
//     sample.kind: R_unif_index by rejection from rbits()).
// fastkpc/src/full_cuda_ci_one_call.cpp:464-567 draws the same tables from
// R::unif_rand() on the host; here the whole stream lives in this library, so
// a caller without R gets the same indices from set.seed semantics and an R
// caller can pass .Random.seed in and take the advanced state back.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace kpc {

class RStream {
 public:
  static constexpr int kN = 624;
  static constexpr int kM = 397;

  // set.seed(seed): 50 rounds of seed = 69069 seed + 1, then one more round
  // per state word (the first of 625 fills the position slot, which FixupSeeds
  // then sets to 624 so the first draw regenerates the block).
  void set_seed(uint32_t seed) {
    for (int j = 0; j < 50; ++j) seed = 69069u * seed + 1u;
    uint32_t words[kN + 1];
    for (int j = 0; j < kN + 1; ++j) {
      seed = 69069u * seed + 1u;
      words[j] = seed;
    }
    std::memcpy(mt_, words + 1, sizeof(mt_));
    mti_ = kN;
  }

  // .Random.seed[2:626] layout: the position, then the 624 state words
  void load(const int32_t* s625) {
    mti_ = s625[0] <= 0 ? kN : s625[0];
    for (int i = 0; i < kN; ++i) mt_[i] = (uint32_t)s625[1 + i];
  }
  void save(int32_t* s625) const {
    s625[0] = mti_;
    for (int i = 0; i < kN; ++i) s625[1 + i] = (int32_t)mt_[i];
  }

  struct State {
    uint32_t mt[kN];
    int mti;
  };
  State state() const {
    State s;
    std::memcpy(s.mt, mt_, sizeof(mt_));
    s.mti = mti_;
    return s;
  }
  void restore(const State& s) {
    std::memcpy(mt_, s.mt, sizeof(mt_));
    mti_ = s.mti;
  }

  // unif_rand() for the Mersenne-Twister kind: the tempered 32-bit word
  // times 2^-32, kept inside (0, 1)
  double unif_rand() {
    const double x = (double)next_word() * 2.3283064365386963e-10;
    const double i2_32m1 = 2.328306437080797e-10;  // 1 / (2^32 - 1)
    if (x <= 0.0) return 0.5 * i2_32m1;
    if (1.0 - x <= 0.0) return 1.0 - 0.5 * i2_32m1;
    return x;
  }

  // R_unif_index(dn), "Rejection" sample.kind: draw ceil(log2(dn)) random
  // bits (16 per uniform) until the value is below dn
  double unif_index(double dn) {
    if (dn <= 0.0) return 0.0;
    const int bits = (int)std::ceil(std::log2(dn));
    double dv;
    do {
      int64_t v = 0;
      for (int k = 0; k <= bits; k += 16) {
        const int v1 = (int)std::floor(unif_rand() * 65536.0);
        v = 65536 * v + v1;
      }
      dv = (double)(v & ((int64_t(1) << bits) - 1));
    } while (dn <= dv);
    return dv;
  }

  // sample(n) (sample.int(n, n) without replacement), 0-based into out
  template <typename T>
  void sample(int n, int* pool, T* out) {
    for (int i = 0; i < n; ++i) pool[i] = i;
    int rem = n;
    for (int i = 0; i < n; ++i) {
      const int j = (int)unif_index((double)rem);
      out[i] = (T)pool[j];
      pool[j] = pool[--rem];
    }
  }

  // energy's permute(J, n): n - 1 swaps of a position drawn below m with the
  // last of the first m positions
  void energy_permute(int n, int* J) {
    int m = n;
    for (int i = 0; i < n - 1; ++i) {
      int j = (int)(m * unif_rand());
      if (j >= m) j = m - 1;
      const int j0 = J[j];
      J[j] = J[--m];
      J[m] = j0;
    }
  }

 private:
  uint32_t next_word() {
    static const uint32_t mag01[2] = {0x0u, 0x9908b0dfu};
    if (mti_ >= kN) {
      int kk;
      uint32_t y;
      for (kk = 0; kk < kN - kM; ++kk) {
        y = (mt_[kk] & 0x80000000u) | (mt_[kk + 1] & 0x7fffffffu);
        mt_[kk] = mt_[kk + kM] ^ (y >> 1) ^ mag01[y & 1u];
      }
      for (; kk < kN - 1; ++kk) {
        y = (mt_[kk] & 0x80000000u) | (mt_[kk + 1] & 0x7fffffffu);
        mt_[kk] = mt_[kk + (kM - kN)] ^ (y >> 1) ^ mag01[y & 1u];
      }
      y = (mt_[kN - 1] & 0x80000000u) | (mt_[0] & 0x7fffffffu);
      mt_[kN - 1] = mt_[kM - 1] ^ (y >> 1) ^ mag01[y & 1u];
      mti_ = 0;
    }
    uint32_t y = mt_[mti_++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }

  uint32_t mt_[kN] = {};
  int mti_ = kN + 1;
};

}  // namespace kpc
