// Process-wide semantic caches of the skeleton path, the counterparts of the
// reference's CompactResultCache and TargetStateCache
// (fastkpc/src/full_cuda_ci_one_call.cpp:1443-1755) and their control
// (:6812-6900):
//   * results: one CI test (dataset, the options that change its value, x, y,
//     S) -> {p, statistic, mean, variance, decision margin}; a replayed
//     skeleton finds every test here and runs none;
//   * target states: one regrXonS fit (dataset, S, target) -> its fit record
//     (selected log sp, GCV score, edf, ..., decision margin); a hit replays
//     the final fit at the cached log sp (same design, same evaluation,
//     bit-identical residuals) instead of the optimiser.
// Both are least-recently-used maps with the reference's default and maximum
// capacities; every entry carries its dataset digest, so per-dataset counts
// and a digest mismatch check are possible.  All members are thread-safe.
#pragma once
#include <algorithm>
#include <array>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <list>
#include <mutex>
#include <unordered_map>

namespace kpc {

struct Digest {
  uint64_t hi = 0, lo = 0;
  bool operator==(const Digest& o) const { return hi == o.hi && lo == o.lo; }
};

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// 128-bit content digest of a column-major n x p matrix of doubles (two
// independent 64-bit chains over the bit patterns, with the shape folded in)
inline Digest dataset_digest(const double* d, int n, int p) {
  Digest g;
  g.hi = mix64(0x6b70636d69ULL ^ (uint64_t)n);
  g.lo = mix64(0x64617461ULL ^ ((uint64_t)p << 32));
  const size_t cnt = (size_t)n * (size_t)p;
  for (size_t i = 0; i < cnt; ++i) {
    uint64_t b;
    std::memcpy(&b, d + i, sizeof(b));
    g.hi = mix64(g.hi ^ b);
    g.lo = mix64((g.lo + 0x632BE59BD9B4E019ULL) ^ (b * 0xD6E8FEB86659FD93ULL));
  }
  return g;
}

struct CacheKey {
  Digest dataset;
  uint64_t a = 0, b = 0, c = 0;
  bool operator==(const CacheKey& o) const {
    return dataset == o.dataset && a == o.a && b == o.b && c == o.c;
  }
};

struct CacheKeyHash {
  size_t operator()(const CacheKey& k) const {
    return (size_t)mix64(k.dataset.hi ^ mix64(k.dataset.lo ^ mix64(k.a ^ mix64(k.b ^ k.c))));
  }
};

struct CacheStats {
  long long capacity = 0, entries = 0, dataset_entries = 0, requests = 0, hits = 0, misses = 0,
            inserts = 0, evictions = 0, generation = 1;
};

template <typename Value>
class LruCache {
 public:
  LruCache(long long capacity, long long max_capacity)
      : capacity_(capacity), max_capacity_(max_capacity) {}

  bool lookup(const CacheKey& k, Value* out) {
    std::lock_guard<std::mutex> lock(mu_);
    ++requests_;
    auto it = map_.find(k);
    if (it == map_.end()) {
      ++misses_;
      return false;
    }
    ++hits_;
    order_.splice(order_.begin(), order_, it->second.pos);
    *out = it->second.value;
    return true;
  }

  void insert(const CacheKey& k, const Value& v) {
    std::lock_guard<std::mutex> lock(mu_);
    auto it = map_.find(k);
    if (it != map_.end()) {
      it->second.value = v;
      order_.splice(order_.begin(), order_, it->second.pos);
      return;
    }
    while ((long long)map_.size() >= capacity_) evict_one();
    order_.push_front(k);
    map_.emplace(k, Entry{v, order_.begin()});
    ++inserts_;
  }

  // capacity in [1, max]; returns false when out of range
  bool configure(long long capacity) {
    if (capacity < 1 || capacity > max_capacity_) return false;
    std::lock_guard<std::mutex> lock(mu_);
    capacity_ = capacity;
    while ((long long)map_.size() > capacity_) evict_one();
    ++generation_;
    return true;
  }

  void reset() {
    std::lock_guard<std::mutex> lock(mu_);
    map_.clear();
    order_.clear();
    requests_ = hits_ = misses_ = inserts_ = evictions_ = 0;
    ++generation_;
  }

  CacheStats stats(const Digest* dataset = nullptr) const {
    std::lock_guard<std::mutex> lock(mu_);
    CacheStats s;
    s.capacity = capacity_;
    s.entries = (long long)map_.size();
    if (dataset)
      for (const auto& kv : map_) s.dataset_entries += kv.first.dataset == *dataset;
    s.requests = requests_;
    s.hits = hits_;
    s.misses = misses_;
    s.inserts = inserts_;
    s.evictions = evictions_;
    s.generation = generation_;
    return s;
  }

 private:
  struct Entry {
    Value value;
    typename std::list<CacheKey>::iterator pos;
  };
  void evict_one() {
    if (order_.empty()) return;
    map_.erase(order_.back());
    order_.pop_back();
    ++evictions_;
  }
  mutable std::mutex mu_;
  long long capacity_, max_capacity_;
  std::list<CacheKey> order_;
  std::unordered_map<CacheKey, Entry, CacheKeyHash> map_;
  long long requests_ = 0, hits_ = 0, misses_ = 0, inserts_ = 0, evictions_ = 0;
  long long generation_ = 1;
};

struct CachedResult {
  double p, stat, mean, var, margin;
};

struct CachedFit {
  std::array<double, 32> record;  // the fit's info row (kMagicInfo doubles)
};

// the reference's capacities (fastkpc/src/full_cuda_ci_one_call.cpp:60-63)
inline LruCache<CachedResult>& result_cache() {
  static LruCache<CachedResult> c(262144, 1048576);
  return c;
}
inline LruCache<CachedFit>& target_cache() {
  static LruCache<CachedFit> c(131072, 524288);
  return c;
}
inline std::atomic<long long>& cache_epoch() {
  static std::atomic<long long> e{1};
  return e;
}

}  // namespace kpc
