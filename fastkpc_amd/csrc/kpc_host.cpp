// Host controller and C ABI of the MI355X kernel-PC CI path.
//
// Skeleton semantics follow pcalg skeleton(method = "stable") as replicated by
// fastkpc/src/skeleton_task_scheduler.cpp:70-132 (task order) and
// fastkpc/src/full_cuda_ci_one_call.cpp:5229-5580 (level loop, per-edge early
// stop, canonical replay, pMax, sepsets, n.edgetests, `done` rule).
//
// CI methods: dcc.gamma (kpcalg/R/dcovgamma.R), hsic.gamma / hsic.perm
// (kpcalg/R/hsicgamma.R, hsicperm.R) and dcc.perm (energy::dcov.test).  Every
// method is symmetric in (x, y), so a test is identified by ({x, y}, S): the
// second orientation of a test already evaluated in the level reuses its
// result, and permutation tests draw their indices from the seed of that
// identity (permuting the variable with the larger index).
//
// Execution (KPC_BATCHED): each level is evaluated in frontier waves.  A wave
// takes, for every undecided edge, the next `chunk` tasks of its current
// orientation (chunk doubles per wave), evaluates all of them in one batch of
// launches -- smooth setups, design geometries, GCV fits, dcov components,
// pair statistics -- and replays them in canonical order.  Device-resident
// caches keyed by S (designs), (target, S) (residuals, components) and by
// column (univariate smooths) make every residual and component computed at
// most once per level.  KPC_PLAIN runs the same kernels once per test with
// no cross-test caching (one launch sequence per CI test).
#include "../../include/kpcmi.h"
#include "kpc_common.h"
#include "kpc_device.h"
#include "kpc_rstream.h"
#include "kpc_cache.h"
#include "kpc_tune.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_last_error;

}  // namespace

namespace kpc {
// kpc_last_error() text for entry points defined in other translation units
void set_last_error(const char* msg) { g_last_error = msg; }
}  // namespace kpc

namespace {

struct KpcError : std::runtime_error {
  int code;
  KpcError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw KpcError(-2, std::string(what) + ": " + hipGetErrorString(e));
}

// Process-wide cache of device blocks, best fit by size.  A released block
// is handed out again only to the thread (and device) that released it: a
// thread's calls run one after another, every stream that used the block is
// either that same stream (stream order protects the reuse) or was
// synchronised before its arenas were released (engine destructor, stream
// guard), so a reused block is never still in use by an earlier kernel.
struct PoolBlock {
  void* ptr;
  size_t bytes;
  int device;
  std::thread::id owner;
};
std::mutex g_pool_mutex;
std::vector<PoolBlock> g_pool;

void pool_empty();

void* pool_acquire(size_t bytes, size_t* got = nullptr) {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  {
    std::lock_guard<std::mutex> lock(g_pool_mutex);
    size_t best = (size_t)-1;
    const std::thread::id me = std::this_thread::get_id();
    for (size_t i = 0; i < g_pool.size(); ++i)
      if (g_pool[i].device == dev && g_pool[i].owner == me && g_pool[i].bytes >= bytes && g_pool[i].bytes <= 4 * bytes + (1 << 20) &&
          (best == (size_t)-1 || g_pool[i].bytes < g_pool[best].bytes))
        best = i;
    if (best != (size_t)-1) {
      void* p = g_pool[best].ptr;
      if (got) *got = g_pool[best].bytes;
      g_pool.erase(g_pool.begin() + (long)best);
      return p;
    }
  }
  // round up so slightly larger later requests reuse the block
  size_t rounded = bytes < (1 << 20) ? bytes : (bytes + (bytes >> 3) + 255) & ~(size_t)255;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, rounded);
  if (e != hipSuccess) {
    // out of memory with cached blocks: give them back and retry once
    pool_empty();
    hip_check(hipMalloc(&p, rounded), "hipMalloc");
  }
  if (got) *got = rounded;
  return p;
}

// Frees every cached block of every thread: blocks are taken out of the pool
// first, then the device drains, so work enqueued before a block's release
// has finished when it is freed.
void pool_empty() {
  std::vector<PoolBlock> drop;
  {
    std::lock_guard<std::mutex> lock(g_pool_mutex);
    drop.swap(g_pool);
  }
  if (drop.empty()) return;
  (void)hipDeviceSynchronize();
  for (const PoolBlock& b : drop) (void)hipFree(b.ptr);
}

void pool_release(void* p, size_t bytes) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(g_pool_mutex);
  g_pool.push_back({p, bytes, dev, std::this_thread::get_id()});
}

// A stream that is synchronised and destroyed when the guard goes out of
// scope.  Declared after the arrays its work uses, it is destroyed first, so
// no kernel is still running when they return their blocks to the pool.
struct StreamGuard {
  hipStream_t s = nullptr;
  StreamGuard() { hip_check(hipStreamCreate(&s), "hipStreamCreate"); }
  StreamGuard(const StreamGuard&) = delete;
  StreamGuard& operator=(const StreamGuard&) = delete;
  ~StreamGuard() {
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
};

template <typename T>
struct DevArray {
  T* ptr = nullptr;
  size_t cap = 0;
  // Blocks come from the process-wide device pool: growing an arena returns
  // the old block to the pool (stream-ordered reuse by the same stream of a
  // later call), so repeated calls allocate nothing once warm.
  void reserve(size_t count) {
    if (count <= cap) return;
    size_t c = count < 64 ? 64 : count;
    size_t bytes = c * sizeof(T);
    size_t got = bytes;
    void* fresh = pool_acquire(bytes, &got);
    if (ptr) pool_release(ptr, cap * sizeof(T));
    ptr = (T*)fresh;
    cap = got / sizeof(T);
  }
  void upload(const std::vector<T>& v, hipStream_t s) {
    reserve(v.size());
    if (!v.empty())
      hip_check(hipMemcpyAsync(ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s),
                "hipMemcpyAsync H2D");
  }
  ~DevArray() {
    if (ptr) pool_release(ptr, cap * sizeof(T));
  }
};

struct KeyHash {
  size_t operator()(const std::pair<int, uint64_t>& k) const {
    uint64_t x = k.second * 0x9E3779B97F4A7C15ULL ^ (uint64_t)(k.first + 1) * 0xC2B2AE3D27D4EB4FULL;
    x ^= x >> 29;
    return (size_t)x;
  }
};

struct TwinHash {
  size_t operator()(const std::pair<uint64_t, uint64_t>& k) const {
    uint64_t x = k.first * 0x9E3779B97F4A7C15ULL ^ (k.second + 0x632BE59BD9B4E019ULL) * 0xC2B2AE3D27D4EB4FULL;
    x ^= x >> 31;
    return (size_t)x;
  }
};

// Open-addressing hash map (linear probing, power-of-two capacity, at most
// half full, Fibonacci hashing of H's value): the per-task lookups of a wave
// (twins, designs, residual and component slots) without a node allocation per
// insert.  Capacity is kept across clear().  No iteration order is exposed.
template <typename K, typename V, typename H>
class FlatMap {
 public:
  V* find(const K& k) {
    if (size_ == 0) return nullptr;
    for (size_t i = slot(k);; i = (i + 1) & mask_) {
      if (!used_[i]) return nullptr;
      if (keys_[i] == k) return &vals_[i];
    }
  }
  // Inserts (k, v) when k is absent; returns the stored value either way.
  V& emplace(const K& k, const V& v) {
    if (2 * (size_ + 1) > used_.size()) grow();
    size_t i = slot(k);
    for (; used_[i]; i = (i + 1) & mask_)
      if (keys_[i] == k) return vals_[i];
    used_[i] = 1;
    keys_[i] = k;
    vals_[i] = v;
    ++size_;
    return vals_[i];
  }
  void set(const K& k, const V& v) { emplace(k, v) = v; }
  void clear() {
    if (size_ == 0) return;
    std::fill(used_.begin(), used_.end(), (unsigned char)0);
    size_ = 0;
  }
  size_t size() const { return size_; }
  // Room for n keys without growing (capacity a power of two >= 2 n).
  void reserve(size_t n) {
    size_t cap = used_.empty() ? 64 : used_.size();
    while (cap < 2 * n) cap *= 2;
    if (cap > used_.size()) rehash(cap);
  }

 private:
  size_t slot(const K& k) const {
    return (size_t)(((uint64_t)H()(k) * 0x9E3779B97F4A7C15ULL) >> shift_);
  }
  void grow() { rehash(used_.empty() ? 64 : 2 * used_.size()); }
  void rehash(size_t cap) {
    std::vector<K> ok;
    std::vector<V> ov;
    std::vector<unsigned char> ou;
    ok.swap(keys_);
    ov.swap(vals_);
    ou.swap(used_);
    keys_.resize(cap);
    vals_.resize(cap);
    used_.assign(cap, 0);
    mask_ = cap - 1;
    shift_ = 64;
    for (size_t c = cap; c > 1; c >>= 1) --shift_;
    size_ = 0;
    for (size_t i = 0; i < ou.size(); ++i)
      if (ou[i]) emplace(ok[i], ov[i]);
  }
  std::vector<K> keys_;
  std::vector<V> vals_;
  std::vector<unsigned char> used_;
  size_t mask_ = 0, size_ = 0;
  int shift_ = 64;
};

struct U64Hash {
  size_t operator()(uint64_t k) const { return (size_t)(k ^ (k >> 31)); }
};

struct TestResult {
  double p, stat, mean, var, margin;
};

// Task seed of the permutation tests: SplitMix64 finaliser folded over (user
// seed, method, min(x,y), max(x,y), |S|, S ascending).  Replaces the
// reference's SHA-256 task identity (fastkpc/src/full_cuda_ci_one_call.cpp:
// 296-325); the oracle computes the same value.
uint64_t splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t task_seed(uint64_t user_seed, int method, int x, int y, const int* S, int s_len) {
  const int a = std::min(x, y), b = std::max(x, y);
  uint64_t h = splitmix(user_seed ^ (uint64_t)(method + 1));
  h = splitmix(h ^ (uint64_t)(a + 1));
  h = splitmix(h ^ (uint64_t)(b + 1));
  h = splitmix(h ^ (uint64_t)(s_len + 1));
  for (int i = 0; i < s_len; ++i) h = splitmix(h ^ (uint64_t)(S[i] + 1));
  return h;
}

struct Task {
  int x, y, slen;
  int S[64];
  uint64_t mask;
  double p = 0.0, stat = 0.0, mean = 0.0, var = 0.0;
  double margin = 1e300;  // smallest decision margin (fits, inchol pivots, permutation ties)
  long long rtab = -1;    // R-stream permutation table (offset in the engine's host tables)
  bool computed = false;
  bool consumed = false;
  bool deleted = false;
};

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Page-locked host block (hipHostMalloc) that grows on demand.  Copies to or
// from it are asynchronous DMA; copies from pageable memory block the host
// while the driver stages them (~15-30 us per work list, ~0.3 ms for a
// level's fit records).  Its owner reuses it only after synchronising the
// streams its earlier copies ran on.
struct PinnedBuf {
  unsigned char* ptr = nullptr;
  size_t cap = 0;
  void reserve(size_t bytes) {
    if (bytes <= cap) return;
    const size_t c = std::max(bytes + bytes / 2, (size_t)1 << 16);
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
    hip_check(hipHostMalloc((void**)&ptr, c, hipHostMallocDefault), "hipHostMalloc");
    cap = c;
  }
};

// HIP stream and timing events of an engine, reused by the later engines of
// the same thread and device (creating and destroying them costs ~0.4 ms per
// call), and the page-locked blocks of its uploads (one per launch lane) and
// read-backs.  They are kept until the process ends; an engine created while
// another engine of the same thread holds them gets a set of its own.
struct EngineHipResources {
  int device = -1;
  bool in_use = false;
  hipStream_t stream = nullptr, stream2 = nullptr;
  hipEvent_t ev[11] = {};
  hipEvent_t ev2[6] = {};
  hipEvent_t evp = nullptr;  // end of the univariate smooths built ahead
  hipEvent_t evq[5] = {};    // stage boundaries of that chain (smooth, geometry, fits, components)
  // third stream: the next level's designs, fits and components built ahead
  // (evs[0..3] bound the geometry, GCV and component launches; evs[3] ends
  // the chain)
  hipStream_t stream3 = nullptr;
  hipEvent_t evs[4] = {};
  PinnedBuf up[2];
  PinnedBuf down;
  PinnedBuf pre;    // their work list
  PinnedBuf ahead;  // the next level's work lists
  PinnedBuf chainc;  // a chain's component lists
};
thread_local std::deque<EngineHipResources>* g_engine_res = nullptr;

EngineHipResources* acquire_engine_resources() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (!g_engine_res) g_engine_res = new std::deque<EngineHipResources>();
  for (auto& r : *g_engine_res)
    if (r.device == dev && !r.in_use) {
      r.in_use = true;
      return &r;
    }
  EngineHipResources r;
  r.device = dev;
  hip_check(hipStreamCreate(&r.stream), "hipStreamCreate");
  hip_check(hipStreamCreate(&r.stream2), "hipStreamCreate");
  for (int e = 0; e < 11; ++e) hip_check(hipEventCreate(&r.ev[e]), "hipEventCreate");
  for (int e = 0; e < 6; ++e) hip_check(hipEventCreate(&r.ev2[e]), "hipEventCreate");
  hip_check(hipEventCreate(&r.evp), "hipEventCreate");
  for (int e = 0; e < 5; ++e) hip_check(hipEventCreate(&r.evq[e]), "hipEventCreate");
  hip_check(hipStreamCreate(&r.stream3), "hipStreamCreate");
  for (int e = 0; e < 4; ++e) hip_check(hipEventCreate(&r.evs[e]), "hipEventCreate");
  r.in_use = true;
  g_engine_res->push_back(r);
  return &g_engine_res->back();
}

uint64_t mask_of(const int* S, int len);
bool next_comb(int* idx, int k, int len);

template <class T>
void swap_arena(DevArray<T>& a, DevArray<T>& b) {
  std::swap(a.ptr, b.ptr);
  std::swap(a.cap, b.cap);
}

// Component arenas (per component slot: k values, n x comp_k vectors, sums,
// moment, status, margin; HSIC permutation rows)
struct CompSet {
  DevArray<double> vals, vecs, dsum, mom, margin, rows;
  DevArray<int> status;
};

class Engine {
 public:
  // skeleton engines index variables with 64-bit sets (p <= 64); the
  // regrXonS / GCV entry points key designs by id and take any p >= 2
  Engine(const double* h_data, int n, int p, const kpc_options& opt, bool skeleton = true)
      : n_(n), p_(p), opt_(opt) {
    static_assert(sizeof(Task::S) / sizeof(int) >= 62, "conditioning sets up to 62");
    if (n < 6) throw KpcError(-1, "n must be > 5");
    if (n > kpc::kMaxN) throw KpcError(-1, "n exceeds the supported maximum (2048)");
    if (p < 2 || (skeleton && p > 64)) throw KpcError(-1, "p must be in [2, 64]");
    if (opt.num_col <= 0 || opt.num_col >= n || opt.num_col > kpc::kMaxNumCol)
      throw KpcError(-1, "num_col must be positive, < n and <= 64");
    if (opt.ci_method < KPC_DCC_GAMMA || opt.ci_method > KPC_HSIC_PERM)
      throw KpcError(-3, "ci_method must be dcc.gamma, hsic.gamma, dcc.perm or hsic.perm");
    method_ = opt.ci_method;
    if ((method_ == KPC_HSIC_GAMMA || method_ == KPC_HSIC_PERM) &&
        !(opt.hsic_sig > 0.0 && std::isfinite(opt.hsic_sig)))
      throw KpcError(-1, "hsic sig must be positive and finite");
    if ((method_ == KPC_DCC_PERM || method_ == KPC_HSIC_PERM) &&
        (opt.replicates < 0 || opt.replicates > 1000000))
      throw KpcError(-1, "replicates must be in [0, 1e6]");
    // energy::dcov.test falls back to index 1 outside [0, 2]
    index_ = (opt.index >= 0.0 && opt.index <= 2.0) ? opt.index : 1.0;
    seed_ = opt.has_seed ? (uint64_t)opt.seed : 0ULL;
    if (opt.perm_stream != KPC_PERM_COUNTER && opt.perm_stream != KPC_PERM_R_STREAM)
      throw KpcError(-1, "perm_stream must be KPC_PERM_COUNTER or KPC_PERM_R_STREAM");
    rmode_ = opt.perm_stream == KPC_PERM_R_STREAM &&
             (method_ == KPC_DCC_PERM || method_ == KPC_HSIC_PERM);
    if (rmode_) {
      if (opt.r_state) rs_.load(opt.r_state);
      else if (opt.has_seed) rs_.set_seed((uint32_t)opt.seed);
      else throw KpcError(-1, "R-stream permutations need a seed or a stream state");
    }
    comp_k_ = method_ == KPC_DCC_PERM ? 2 : opt.num_col;
    for (long long i = 0; i < (long long)n * p; ++i)
      if (!std::isfinite(h_data[i])) throw KpcError(-1, "data must be finite");
    if (opt.use_cache) {
      // a test's value depends on the data, the method and that method's
      // own parameters; a fit's only on the data, S and the target
      dataset_ = kpc::dataset_digest(h_data, n, p);
      use_targets_ = true;
      use_results_ = !rmode_;
      uint64_t h = kpc::mix64(0x726573756c74ULL ^ (uint64_t)method_);
      auto fold = [&h](uint64_t v) { h = kpc::mix64(h ^ v); };
      auto bits = [](double v) {
        uint64_t b;
        std::memcpy(&b, &v, sizeof(b));
        return b;
      };
      if (method_ != KPC_DCC_PERM) fold((uint64_t)opt.num_col);
      if (method_ == KPC_HSIC_GAMMA || method_ == KPC_HSIC_PERM) fold(bits(opt.hsic_sig));
      if (method_ == KPC_DCC_PERM) fold(bits(index_));
      if (method_ == KPC_DCC_PERM || method_ == KPC_HSIC_PERM) {
        fold((uint64_t)opt.replicates);
        fold((uint64_t)opt.include_observed);
        fold(seed_);
      }
      opt_hash_ = h;
    }
    data_bytes_ = sizeof(double) * (size_t)n * p;
    d_data_ = (double*)pool_acquire(data_bytes_);
    hip_check(hipMemcpy(d_data_, h_data, sizeof(double) * (size_t)n * p, hipMemcpyHostToDevice),
              "upload data");
    k_ = opt.num_col;
    ctl_.tol = 1e-7;
    ctl_.max_half_single = 15;
    ctl_.max_half_multi = 25;
    ctl_.max_iterations = 400;
    ctl_.max_newton_step = 5.0;
    ctl_.probe_step = 2.0;
    ctl_.max_probes = 5;
    ctl_.rank_tol = 1.4901161193847656e-8;

    if (const char* sh = std::getenv("KPC_DCC_SHAPE")) {
      int jt = 0, rb = 0;
      if (std::sscanf(sh, "%dx%d", &jt, &rb) == 2 && kpc::dcc_perm_shape_supported(jt, rb) &&
          rb > 1) {
        dcc_perm_jt_ = jt;
        dcc_perm_rb_ = rb;
        dcc_shape_forced_ = true;
      }
    }
    res_ = acquire_engine_resources();
    stream_ = res_->stream;
    stream2_ = res_->stream2;
    stream3_ = res_->stream3;
    for (int e = 0; e < 8; ++e) ev_[e] = res_->ev[e];
    for (int e = 0; e < 3; ++e) tev_[e] = res_->ev[8 + e];
    for (int e = 0; e < 6; ++e) ev2_[e] = res_->ev2[e];
    if (const char* env = std::getenv("KPC_LANES")) lanes_ = std::atoi(env) == 1 ? 1 : 2;
    if (const char* env = std::getenv("KPC_AHEAD_COMPONENTS")) comps_ahead_ = std::atoi(env) != 0;
  }
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;
  ~Engine() {
    // arenas go back to the pool after this body: nothing may still use them
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (stream2_) (void)hipStreamSynchronize(stream2_);
    if (stream3_) (void)hipStreamSynchronize(stream3_);
    if (d_data_) pool_release(d_data_, data_bytes_);
    if (res_) res_->in_use = false;  // stream and events stay with the thread
  }

  // ---- R's global stream (KPC_PERM_R_STREAM) --------------------------
  bool r_stream() const { return rmode_; }
  kpc::RStream::State rstate() const { return rs_.state(); }
  void set_rstate(const kpc::RStream::State& s) { rs_.restore(s); }
  void save_rstate(int* out) const { rs_.save(out); }
  void clear_tables() { rtab_host_.clear(); }
  // the test's R x n index table, drawn now from the stream: dcc.perm as
  // energy dcov.test (one index vector mutated by every replicate),
  // hsic.perm as sample(n) per replicate
  void draw_table(Task& t) {
    const int R = opt_.replicates;
    t.rtab = (long long)rtab_host_.size();
    rtab_host_.resize(rtab_host_.size() + (size_t)R * n_);
    unsigned short* out = rtab_host_.data() + t.rtab;
    pool_.resize((size_t)n_);
    if (method_ == KPC_DCC_PERM) {
      std::vector<int>& J = pool_;
      for (int i = 0; i < n_; ++i) J[i] = i;
      for (int r = 0; r < R; ++r) {
        rs_.energy_permute(n_, J.data());
        for (int i = 0; i < n_; ++i) out[(size_t)r * n_ + i] = (unsigned short)J[i];
      }
    } else {
      for (int r = 0; r < R; ++r) rs_.sample(n_, pool_.data(), out + (size_t)r * n_);
    }
  }

  // ---- per-level arenas / caches -------------------------------------
  void begin_level(int level) {
    level_ = level;
    l1_claim_ = false;
    ahead_claim_ = false;
    claim_comps_ = false;
    design_map_.clear();
    resid_map_.clear();
    comp_map_.clear();
    twin_map_.clear();
    n_design_ = n_resid_ = n_comp_ = n_lsmooth_ = 0;
    if (level <= 2) {
      m_ = 1;
      dim_ = level == 2 ? 29 : 9;
      rk_ = level == 2 ? 27 : 8;
      q_ = level == 2 ? 30 : 10;
    } else {
      m_ = level;
      dim_ = 9;
      rk_ = 8;
      q_ = 1 + 9 * level;
    }
    if (level > 0) {
      if (q_ >= n_) throw KpcError(-4, "model coefficient dimension must be less than n");
      if (q_ > kpc::kMaxQ || m_ > kpc::kMaxPen)
        throw KpcError(-4, "conditioning set too large (|S| <= 62)");
      if (m_ * rk_ > q_) throw KpcError(-4, "penalty layout exceeds the GCV scratch bound");
      layout_ = kpc::host_design_layout(n_, q_, m_, rk_);
    }
  }

  // Univariate smooths (levels 1 and >= 3, the same set-up for a column at
  // either level) go to the per-column cache, built once per call; level 1
  // keeps per-design slots when its tests run one launch sequence each (the
  // plain form).
  void set_share_univariate(bool v) { share_uni_l1_ = v; }
  bool uni_smooths() const { return level_ > 2 || (level_ == 1 && share_uni_l1_); }

  // Built ahead on lane 1 while level 0 runs on lane 0 (all of it depends on
  // the data alone): every column's univariate smooth and, outside the R
  // stream, the caches and graph capture, level 1's designs (S = {s}, design
  // slot s) and residual fits ((t, {s}), residual slot t p + s).  Level 1
  // claims what its tests use (ensure_design, evaluate): only claimed work is
  // checked (a failed set-up or fit is an error only where a test uses it) and
  // counted, and the first wave that claims anything orders lane 0 after this
  // chain.  Work lists, slabs and the level-1 arenas are sized here once;
  // level 1's own reservations never exceed them.
  void prefetch_ahead() {
    if (uni_slot_.empty()) uni_slot_.assign(p_, -1);
    uni_ready_.assign(p_, 0);
    std::vector<int> items;
    for (int c = 0; c < p_; ++c)
      if (uni_slot_[c] < 0) {
        items.insert(items.end(), {1, c, c, c});
        uni_ready_[c] = 1;
      }
    if (items.empty()) return;
    const int q1 = 10, m1 = 1, dim1 = 9, rk1 = 8;
    const bool with_level1 =
        !rmode_ && !use_targets_ && !use_results_ && !use_graph_ && n_ > q1 && (int)items.size() == 4 * p_;
    std::vector<kpc::GeometryItem> geo;
    std::vector<kpc::MagicItem> fits;
    if (with_level1) {
      for (int c = 0; c < p_; ++c) {
        kpc::GeometryItem gi;
        std::memset(&gi, 0, sizeof(gi));
        gi.design_slot = c;
        gi.smooth_slot[0] = c;
        geo.push_back(gi);
      }
      for (int t = 0; t < p_; ++t)
        for (int c = 0; c < p_; ++c)
          if (t != c) fits.push_back(kpc::MagicItem{c, t, t * p_ + c, 0});
    }
    const size_t nx = (size_t)n_ * 29;
    usmooth_X_.reserve((size_t)p_ * nx);
    usmooth_S_.reserve((size_t)p_ * 29 * 29);
    usmooth_info_.reserve((size_t)p_ * 4);
    // one page-locked block: the smooth, design and fit work lists
    const size_t b0 = items.size() * sizeof(int), b1 = geo.size() * sizeof(kpc::GeometryItem),
                 b2 = fits.size() * sizeof(kpc::MagicItem);
    const size_t o1 = aligned256(b0), o2 = o1 + aligned256(b1);
    PinnedBuf& pb = res_->pre;
    pb.reserve(o2 + aligned256(b2));
    std::memcpy(pb.ptr, items.data(), b0);
    pre_items_.reserve(items.size());
    hip_check(hipMemcpyAsync(pre_items_.ptr, pb.ptr, b0, hipMemcpyHostToDevice, stream2_),
              "hipMemcpyAsync H2D");
    const int nsm = (int)items.size() / 4;
    const long long stride = kpc::smooth_scratch_doubles(n_);
    const int grid = std::min(nsm, grid_limit(stride, 512));
    pre_scratch_.reserve((size_t)grid * stride);
    hip_check(hipEventRecord(res_->evq[0], stream2_), "event");
    hip_check(kpc::launch_smooth_setup(d_data_, n_, p_, pre_items_.ptr, nsm, usmooth_X_.ptr,
                                       usmooth_S_.ptr, usmooth_info_.ptr, pre_scratch_.ptr, grid,
                                       stream2_),
              "smooth setup launch");
    hip_check(hipEventRecord(res_->evq[1], stream2_), "event");
    if (with_level1) {
      const kpc::DesignLayout lay = kpc::host_design_layout(n_, q1, m1, rk1);
      design_.reserve((size_t)p_ * (size_t)lay.stride);
      design_status_.reserve(p_);
      resid_.reserve((size_t)p_ * p_ * n_);
      rinfo_.reserve((size_t)p_ * p_ * kpc::kMagicInfo);
      std::memcpy(pb.ptr + o1, geo.data(), b1);
      std::memcpy(pb.ptr + o2, fits.data(), b2);
      pre_geo_items_.reserve(geo.size());
      pre_fit_items_.reserve(fits.size());
      hip_check(hipMemcpyAsync(pre_geo_items_.ptr, pb.ptr + o1, b1, hipMemcpyHostToDevice, stream2_),
                "hipMemcpyAsync H2D");
      hip_check(hipMemcpyAsync(pre_fit_items_.ptr, pb.ptr + o2, b2, hipMemcpyHostToDevice, stream2_),
                "hipMemcpyAsync H2D");
      const long long gstride = kpc::geometry_scratch_doubles(q1);
      const int ggrid = std::min((int)geo.size(), grid_limit(gstride, 2048));
      pre_scratch2_.reserve((size_t)ggrid * gstride);
      hip_check(kpc::launch_geometry(pre_geo_items_.ptr, (int)geo.size(), n_, q1, m1, dim1, rk1,
                                     usmooth_X_.ptr, usmooth_S_.ptr, design_.ptr,
                                     design_status_.ptr, (long long)n_ * 29, pre_scratch2_.ptr,
                                     ggrid, stream2_),
                "geometry launch");
      hip_check(hipEventRecord(res_->evq[2], stream2_), "event");
      const long long mstride = kpc::magic_scratch_doubles(q1, m1, rk1, n_);
      const int mgrid = std::min((int)fits.size(), grid_limit(mstride, 65535));
      pre_scratch3_.reserve((size_t)mgrid * mstride);
      hip_check(kpc::launch_magic(pre_fit_items_.ptr, (int)fits.size(), n_, q1, m1, rk1, d_data_,
                                  design_.ptr, resid_.ptr, rinfo_.ptr, pre_scratch3_.ptr, mgrid,
                                  ctl_, stream2_, nullptr),
                "magic launch");
      hip_check(hipEventRecord(res_->evq[3], stream2_), "event");
      l1_ahead_ = true;
      // their components at component slot = residual slot (t p + s), into
      // the chain's component arenas (level 0 uses the engine's)
      l1_comps_ = launch_chain_components(stream2_, fits, resid_.ptr, (size_t)2 * p_ * p_);
      hip_check(hipEventRecord(res_->evq[4], stream2_), "event");
    }
    hip_check(hipEventRecord(res_->evp, stream2_), "event");
    prefetch_pending_ = true;
    prefetch_time_pending_ = true;
  }

  // Level 1 with its designs and fits built ahead: slots are the fixed ones
  // of prefetch_ahead (design s, residual t p + s), none is allocated.
  void adopt_level1() {
    if (!l1_ahead_ || level_ != 1) return;
    l1_claim_ = true;
    n_design_ = p_;
    n_resid_ = p_ * p_;
    if (l1_comps_ && comps_hold(chain_comp_, (size_t)2 * p_ * p_)) {
      swap_comps(chain_comp_);
      n_comp_ = p_ * p_;
      claim_comps_ = true;
    }
  }

  // Level `next`'s additive designs and residual fits, built ahead on the
  // third stream while the current level runs.  Adjacency only shrinks during
  // a level, so every (target, S) that level `next` can fit comes from an edge
  // x - y of the level-start graph `snap` with S a `next`-subset of x's other
  // neighbours there; that superset is built when it has at most `limit` fits
  // and every column's univariate smooth exists.  Level `next` claims what its
  // tests use (adopt_ahead); unclaimed work is neither checked nor counted.
  // The lists are launched behind the first wave's launches (launch_ahead).
  void plan_ahead(const unsigned char* snap, int next, size_t limit) {
    ahead_queued_ = false;
    ahead_level_ = -1;
    if (next < 3 || limit == 0 || rmode_ || use_targets_ || use_results_ || use_graph_) return;
    const int q = 1 + 9 * next;
    if (q >= n_ || q > kpc::kMaxQ || next > kpc::kMaxPen || next * 8 > q) return;
    if (uni_slot_.size() != (size_t)p_ || uni_ready_.size() != (size_t)p_) return;
    for (int c = 0; c < p_; ++c)
      if (uni_slot_[c] < 0 && !uni_ready_[c]) return;  // a smooth not built
    const int p = p_;
    double count = 0.0;  // orientations x subsets (each gives at most two fits)
    for (int a = 0; a < p; ++a) {
      int deg = 0;
      for (int i = 0; i < p; ++i) deg += i != a && snap[(size_t)i * p + a];
      if (deg - 1 < next) continue;
      double c = 1.0;
      for (int i = 0; i < next; ++i) c = c * (double)(deg - 1 - i) / (double)(i + 1);
      count += deg * c;
    }
    if (count == 0.0 || count > 4.0 * (double)limit) return;
    ahead_geo_.clear();
    ahead_fits_.clear();
    ahead_design_map_.clear();
    ahead_resid_map_.clear();
    ahead_design_map_.reserve((size_t)count);
    ahead_resid_map_.reserve(std::min((size_t)(2.0 * count), limit + 2));
    int nb[64], idx[64], S[64];
    for (int x = 0; x < p; ++x)
      for (int y = 0; y < p; ++y) {
        if (x == y || !snap[(size_t)y * p + x]) continue;
        int len = 0;
        for (int i = 0; i < p; ++i)
          if (i != y && snap[(size_t)i * p + x]) nb[len++] = i;
        if (len < next) continue;
        for (int i = 0; i < next; ++i) idx[i] = i;
        do {
          for (int i = 0; i < next; ++i) S[i] = nb[idx[i]];
          const uint64_t mask = mask_of(S, next);
          int ds;
          if (const int* d = ahead_design_map_.find(mask)) {
            ds = *d;
          } else {
            ds = (int)ahead_geo_.size();
            ahead_design_map_.set(mask, ds);
            kpc::GeometryItem gi;
            std::memset(&gi, 0, sizeof(gi));
            gi.design_slot = ds;
            for (int i = 0; i < next; ++i) gi.smooth_slot[i] = S[i];
            ahead_geo_.push_back(gi);
          }
          for (int side = 0; side < 2; ++side) {
            const auto key = std::make_pair(side ? y : x, mask);
            if (ahead_resid_map_.find(key)) continue;
            const int rs = (int)ahead_fits_.size();
            if ((size_t)rs >= limit) return;
            ahead_resid_map_.set(key, rs);
            ahead_fits_.push_back(kpc::MagicItem{ds, key.first, rs, 0});
          }
        } while (next_comb(idx, next, len));
      }
    ahead_level_ = next;
    ahead_queued_ = true;
  }

  // Component arenas of a chain built ahead (the engine's own serve the
  // running level); swapped in by the level that claims them.
  void reserve_comps(CompSet& c, size_t slots) {
    c.vals.reserve(slots * k_);
    c.vecs.reserve(slots * (size_t)n_ * comp_k_);
    c.dsum.reserve(slots);
    c.mom.reserve(slots);
    c.status.reserve(slots);
    c.margin.reserve(slots);
    if (method_ == KPC_HSIC_PERM) c.rows.reserve(slots * (size_t)n_ * 16 * hsic_kt());
  }
  bool comps_hold(const CompSet& c, size_t slots) const {
    return c.vals.cap >= slots * k_ && c.vecs.cap >= slots * (size_t)n_ * comp_k_ &&
           c.dsum.cap >= slots && c.mom.cap >= slots && c.status.cap >= slots &&
           c.margin.cap >= slots &&
           (method_ != KPC_HSIC_PERM || c.rows.cap >= slots * (size_t)n_ * 16 * hsic_kt());
  }
  void swap_comps(CompSet& c) {
    swap_arena(comp_vals_, c.vals);
    swap_arena(comp_vecs_, c.vecs);
    swap_arena(comp_dsum_, c.dsum);
    swap_arena(comp_mom_, c.mom);
    swap_arena(comp_status_, c.status);
    swap_arena(comp_margin_, c.margin);
    swap_arena(comp_rows_, c.rows);
  }
  // The components of a chain's fits (component slot = residual slot), read
  // from `src`, into chain_comp_ with room for `slots`; returns whether they
  // were launched (KPC_AHEAD_COMPONENTS=0 leaves them to the claiming level).
  bool launch_chain_components(hipStream_t stream, const std::vector<kpc::MagicItem>& fits,
                               const double* src, size_t slots) {
    chain_comps_ = false;
    if (!comps_ahead_ || fits.empty()) return false;
    std::vector<long long> csrc(fits.size());
    std::vector<int> cdst(fits.size());
    for (size_t i = 0; i < fits.size(); ++i) {
      csrc[i] = (long long)fits[i].resid_slot * n_;
      cdst[i] = fits[i].resid_slot;
    }
    reserve_comps(chain_comp_, slots);
    const size_t b1 = csrc.size() * sizeof(long long), b2 = cdst.size() * sizeof(int);
    PinnedBuf& pb = res_->chainc;
    pb.reserve(aligned256(b1) + aligned256(b2));
    std::memcpy(pb.ptr, csrc.data(), b1);
    std::memcpy(pb.ptr + aligned256(b1), cdst.data(), b2);
    chain_comp_src_.reserve(csrc.size());
    chain_comp_dst_.reserve(cdst.size());
    hip_check(hipMemcpyAsync(chain_comp_src_.ptr, pb.ptr, b1, hipMemcpyHostToDevice, stream),
              "hipMemcpyAsync H2D");
    hip_check(hipMemcpyAsync(chain_comp_dst_.ptr, pb.ptr + aligned256(b1), b2, hipMemcpyHostToDevice,
                             stream),
              "hipMemcpyAsync H2D");
    const Lane L{stream,           nullptr, nullptr, nullptr, nullptr, &chain_comp_src_,
                 &chain_comp_dst_, nullptr, nullptr, nullptr, nullptr, nullptr,
                 nullptr,          &chain_scratch4_, nullptr};
    swap_comps(chain_comp_);
    launch_components(L, src, fits.size());
    swap_comps(chain_comp_);
    chain_comps_ = true;
    return true;
  }

  // stage times of the chain built during level 0, counted with level 1's
  // first wave (the chain has ended by then: lane 1 ran it)
  void account_prefetch() {
    if (!prefetch_time_pending_ || level_ < 1) return;
    prefetch_time_pending_ = false;
    hip_check(hipEventSynchronize(res_->evp), "event synchronize");
    float ms;
    hip_check(hipEventElapsedTime(&ms, res_->evq[0], res_->evq[1]), "event elapsed");
    stage_[0] += ms * 1e-3;
    ++launches_[0];
    if (l1_ahead_) {
      hip_check(hipEventElapsedTime(&ms, res_->evq[1], res_->evq[2]), "event elapsed");
      stage_[1] += ms * 1e-3;
      hip_check(hipEventElapsedTime(&ms, res_->evq[2], res_->evq[3]), "event elapsed");
      stage_[2] += ms * 1e-3;
      ++launches_[1];
      ++launches_[2];
      if (l1_comps_) {
        hip_check(hipEventElapsedTime(&ms, res_->evq[3], res_->evq[4]), "event elapsed");
        stage_[3] += ms * 1e-3;
        ++launches_[3];
      }
    }
  }

  // stage times of an adopted chain (its events are re-recorded by the next)
  void account_ahead() {
    if (!ahead_time_pending_) return;
    ahead_time_pending_ = false;
    hip_check(hipEventSynchronize(res_->evs[3]), "event synchronize");
    float ms;
    hip_check(hipEventElapsedTime(&ms, res_->evs[0], res_->evs[1]), "event elapsed");
    stage_[1] += ms * 1e-3;
    hip_check(hipEventElapsedTime(&ms, res_->evs[1], res_->evs[2]), "event elapsed");
    stage_[2] += ms * 1e-3;
    ++launches_[1];
    ++launches_[2];
    if (chain_comps_) {
      hip_check(hipEventElapsedTime(&ms, res_->evs[2], res_->evs[3]), "event elapsed");
      stage_[3] += ms * 1e-3;
      ++launches_[3];
    }
  }

  void launch_ahead() {
    ahead_queued_ = false;
    account_ahead();
    // the previous chains' lists, arenas and slabs are reused below
    if (ahead_launched_) hip_check(hipEventSynchronize(res_->evs[3]), "event synchronize");
    if (prefetch_pending_) hip_check(hipEventSynchronize(res_->evp), "event synchronize");
    const int next = ahead_level_;
    const int q = 1 + 9 * next, m = next, dim = 9, rk = 8;
    const kpc::DesignLayout lay = kpc::host_design_layout(n_, q, m, rk);
    const size_t nd = ahead_geo_.size(), nf = ahead_fits_.size();
    // room for as many of the level's own designs and fits again
    ahead_design_.reserve(2 * nd * (size_t)lay.stride);
    ahead_design_status_.reserve(2 * nd);
    ahead_resid_.reserve(2 * nf * (size_t)n_);
    ahead_rinfo_.reserve(2 * nf * kpc::kMagicInfo);
    const size_t b1 = nd * sizeof(kpc::GeometryItem), b2 = nf * sizeof(kpc::MagicItem);
    const size_t o2 = aligned256(b1);
    PinnedBuf& pb = res_->ahead;
    pb.reserve(o2 + aligned256(b2));
    std::memcpy(pb.ptr, ahead_geo_.data(), b1);
    std::memcpy(pb.ptr + o2, ahead_fits_.data(), b2);
    ahead_geo_items_.reserve(nd);
    ahead_fit_items_.reserve(nf);
    // the univariate smooths built ahead on lane 1 (the others were built by
    // waves already synchronised)
    hip_check(hipStreamWaitEvent(stream3_, res_->evp, 0), "stream wait");
    hip_check(hipMemcpyAsync(ahead_geo_items_.ptr, pb.ptr, b1, hipMemcpyHostToDevice, stream3_),
              "hipMemcpyAsync H2D");
    hip_check(hipMemcpyAsync(ahead_fit_items_.ptr, pb.ptr + o2, b2, hipMemcpyHostToDevice, stream3_),
              "hipMemcpyAsync H2D");
    hip_check(hipEventRecord(res_->evs[0], stream3_), "event");
    const long long gstride = kpc::geometry_scratch_doubles(q);
    const int ggrid = std::min((int)nd, grid_limit(gstride, 2048));
    ahead_scratch2_.reserve((size_t)ggrid * gstride);
    hip_check(kpc::launch_geometry(ahead_geo_items_.ptr, (int)nd, n_, q, m, dim, rk, usmooth_X_.ptr,
                                   usmooth_S_.ptr, ahead_design_.ptr, ahead_design_status_.ptr,
                                   (long long)n_ * 29, ahead_scratch2_.ptr, ggrid, stream3_),
              "geometry launch");
    hip_check(hipEventRecord(res_->evs[1], stream3_), "event");
    const long long mstride = kpc::magic_scratch_doubles(q, m, rk, n_);
    const int mgrid = std::min((int)nf, grid_limit(mstride, 65535));
    ahead_scratch3_.reserve((size_t)mgrid * mstride);
    hip_check(kpc::launch_magic(ahead_fit_items_.ptr, (int)nf, n_, q, m, rk, d_data_, ahead_design_.ptr,
                                ahead_resid_.ptr, ahead_rinfo_.ptr, ahead_scratch3_.ptr, mgrid, ctl_,
                                stream3_, nullptr),
              "magic launch");
    hip_check(hipEventRecord(res_->evs[2], stream3_), "event");
    launch_chain_components(stream3_, ahead_fits_, ahead_resid_.ptr, 2 * nf);
    hip_check(hipEventRecord(res_->evs[3], stream3_), "event");
    ahead_launched_ = true;
    ahead_ready_ = true;
  }

  bool claims_ahead() const { return ahead_claim_; }

  // Level `next` (this level) takes the arenas built ahead for it: design
  // and residual slots of the superset first, the level's own after them.
  void adopt_ahead(size_t max_designs, size_t max_resid) {
    const bool ready = ahead_ready_ && ahead_level_ == level_ && level_ > 0;
    ahead_ready_ = false;
    if (!ready) return;
    const size_t nd = ahead_geo_.size(), nf = ahead_fits_.size();
    if ((nd + max_designs) * (size_t)layout_.stride > ahead_design_.cap ||
        nd + max_designs > ahead_design_status_.cap || (nf + max_resid) * (size_t)n_ > ahead_resid_.cap ||
        (nf + max_resid) * (size_t)kpc::kMagicInfo > ahead_rinfo_.cap)
      return;
    const bool with_comps = chain_comps_ && comps_hold(chain_comp_, nf + max_resid);
    swap_arena(design_, ahead_design_);
    swap_arena(design_status_, ahead_design_status_);
    if (with_comps) {
      swap_comps(chain_comp_);
      n_comp_ = (int)nf;
      claim_comps_ = true;
    }
    swap_arena(resid_, ahead_resid_);
    swap_arena(rinfo_, ahead_rinfo_);
    // the maps move too: the lists are refilled for the level after this one
    std::swap(claim_design_map_, ahead_design_map_);
    std::swap(claim_resid_map_, ahead_resid_map_);
    n_design_ = (int)nd;
    n_resid_ = (int)nf;
    ahead_claim_ = true;
    ahead_wait_ = true;
    ahead_time_pending_ = true;
  }

  void reserve_level(size_t max_designs, size_t max_resid) {
    design_map_.reserve(max_designs);
    resid_map_.reserve(max_resid);
    comp_map_.reserve(max_resid);
    const size_t nx = (size_t)n_ * 29;
    if (level_ > 0 && level_ <= 2) {
      lsmooth_X_.reserve(max_designs * nx);
      lsmooth_S_.reserve(max_designs * 29 * 29);
      lsmooth_info_.reserve(max_designs * 4);
    }
    if (level_ == 1 || level_ > 2) {
      usmooth_X_.reserve((size_t)p_ * nx);
      usmooth_S_.reserve((size_t)p_ * 29 * 29);
      usmooth_info_.reserve((size_t)p_ * 4);
    }
    if (level_ > 0) {
      design_.reserve(max_designs * (size_t)layout_.stride);
      design_status_.reserve(max_designs);
      resid_.reserve(max_resid * (size_t)n_);
      rinfo_.reserve(max_resid * kpc::kMagicInfo);
    }
    comp_vals_.reserve(max_resid * (size_t)k_);
    comp_vecs_.reserve(max_resid * (size_t)n_ * comp_k_);
    comp_dsum_.reserve(max_resid);
    comp_mom_.reserve(max_resid);
    comp_status_.reserve(max_resid);
    comp_margin_.reserve(max_resid);
    if (method_ == KPC_HSIC_PERM) comp_rows_.reserve(max_resid * (size_t)n_ * 16 * hsic_kt());
  }

  // Evaluate p-values for the given tasks (all with |S| == level_).
  // with_pairs = false: fits and components only (the tests' statistics are
  // computed by a later call)
  void evaluate(std::vector<Task*>& all, bool with_pairs = true) {
    if (all.empty()) return;
    const double t0 = now_s();
    // orientation-free identity: reuse a twin's result, evaluate each once
    // (not for R-stream tests: each draws its own indices and permutes y)
    std::vector<Task*> tasks;
    std::vector<std::pair<Task*, size_t>> dups;  // (task, index into tasks)
    in_batch_.clear();
    in_batch_.reserve(all.size());
    twin_map_.reserve(twin_map_.size() + all.size());
    wave_claimed_ = false;
    claimed_geos_.clear();
    claimed_fits_.clear();
    claimed_comps_.clear();
    pair_rtab_.clear();
    for (Task* t : all) {
      if (use_results_ && with_pairs) {
        kpc::CachedResult cr;
        if (kpc::result_cache().lookup(result_key(*t), &cr)) {
          apply_result(*t, TestResult{cr.p, cr.stat, cr.mean, cr.var, cr.margin});
          ++result_hits_;
          continue;
        }
      }
      if (rmode_ || !with_pairs) {
        if (with_pairs) pair_rtab_.push_back(t->rtab);
        tasks.push_back(t);
        continue;
      }
      const auto key = twin_key(*t);
      if (const TestResult* done = twin_map_.find(key)) {
        apply_result(*t, *done);
        continue;
      }
      if (const size_t* it = in_batch_.find(key)) {
        dups.emplace_back(t, *it);
        continue;
      }
      in_batch_.emplace(key, tasks.size());
      tasks.push_back(t);
    }
    if (tasks.empty()) return;
    const bool perm = method_ == KPC_DCC_PERM || method_ == KPC_HSIC_PERM;
    // KPC_GRAPH=1: the wave's device pipeline (smooth setups .. pair
    // statistics) is captured into one hipGraph and launched as a unit; the
    // permutation methods synchronise between replicate chunks and stay eager
    const bool graph = use_graph_ && !perm;
    // distinct conditioning sets alternate between the two lanes (waves of
    // fewer than 64 tests stay on one: a second lane only adds launches there)
    std::vector<int> lane_of(tasks.size(), 0);
    int nlanes = 1;
    if (lanes_ > 1 && !graph && !perm && level_ > 0 && tasks.size() >= 64) {
      lane_map_.clear();
      lane_map_.reserve(tasks.size());
      int distinct = 0;
      for (size_t i = 0; i < tasks.size(); ++i) {
        if (const int* l = lane_map_.find(tasks[i]->mask)) {
          lane_of[i] = *l;
          continue;
        }
        lane_of[i] = distinct & 1;
        lane_map_.set(tasks[i]->mask, distinct & 1);
        ++distinct;
      }
      if (distinct > 1) nlanes = 2;
    }
    struct LaneWork {
      std::vector<int> smooth_items;  // d, c0, c1, slot
      std::vector<kpc::GeometryItem> geo_items;
      std::vector<kpc::MagicItem> fit_items;
      std::vector<uint64_t> fit_mask;  // conditioning set of each fit (target-state cache)
      std::vector<double> fixed;       // cached fit records the fits with pad > 0 replay
      std::vector<long long> comp_src;
      std::vector<int> comp_dst;
      std::vector<int> pair_x, pair_y;
      std::vector<int> fit_x, fit_y;  // residual slot of each task's fits (-1: raw column)
      std::vector<size_t> task_idx;
      int nsm = 0;
    };
    LaneWork W[2];
    // above level 2 the smooths are univariate and shared between sets: they
    // run on lane 0 first and lane 1 waits for them
    const bool uni = uni_smooths();
    std::vector<int> shared_smooth;
    std::vector<unsigned long long> seeds;
    for (size_t ti = 0; ti < tasks.size(); ++ti) {
      Task* t = tasks[ti];
      LaneWork& w = W[lane_of[ti]];
      std::vector<int>& smooth_items = uni ? shared_smooth : w.smooth_items;
      int slots[2];
      int rslot[2] = {-1, -1};
      if (perm) seeds.push_back(task_seed(seed_, method_, t->x, t->y, t->S, t->slen));
      for (int side = 0; side < 2; ++side) {
        // side 0: min(x, y), side 1: max(x, y) (the permuted variable); R
        // stream: x, then y (kpcalg permutes the test's y)
        const int tgt = rmode_ ? (side ? t->y : t->x)
                               : (side ? std::max(t->x, t->y) : std::min(t->x, t->y));
        auto key = std::make_pair(tgt, t->mask);
        if (const int* it = comp_map_.find(key)) {
          slots[side] = *it;
          if (const int* rt = resid_map_.find(key)) rslot[side] = *rt;
          continue;
        }
        if (claim_comps_) {
          int rs = -1;
          if (ahead_claim_) {
            if (const int* ahead = claim_resid_map_.find(key)) rs = *ahead;
          } else if (l1_claim_) {
            rs = tgt * p_ + t->S[0];
          }
          if (rs >= 0) {
            // fit and component built ahead (component slot = residual slot)
            const int dslot = ensure_design(*t, smooth_items, w.geo_items);
            comp_map_.set(key, rs);
            resid_map_.set(key, rs);
            slots[side] = rs;
            rslot[side] = rs;
            claimed_fits_.push_back(kpc::MagicItem{dslot, tgt, rs, 0});
            claimed_comps_.push_back(rs);
            if (l1_claim_) wave_claimed_ = true;
            continue;
          }
        }
        const int cs = n_comp_++;
        comp_map_.set(key, cs);
        slots[side] = cs;
        w.comp_dst.push_back(cs);
        if (level_ == 0) {
          w.comp_src.push_back((long long)tgt * n_);
          continue;
        }
        int dslot = ensure_design(*t, smooth_items, w.geo_items);
        if (ahead_claim_) {
          if (const int* ahead = claim_resid_map_.find(key)) {
            // fit built ahead during the previous level
            const int rs = *ahead;
            resid_map_.set(key, rs);
            rslot[side] = rs;
            claimed_fits_.push_back(kpc::MagicItem{dslot, tgt, rs, 0});
            w.comp_src.push_back((long long)rs * n_);
            continue;
          }
        }
        if (l1_claim_) {
          // fit built ahead at residual slot t p + s
          const int rs = tgt * p_ + t->S[0];
          resid_map_.set(key, rs);
          rslot[side] = rs;
          claimed_fits_.push_back(kpc::MagicItem{dslot, tgt, rs, 0});
          wave_claimed_ = true;
          w.comp_src.push_back((long long)rs * n_);
          continue;
        }
        const int rs = n_resid_++;
        resid_map_.set(key, rs);
        rslot[side] = rs;
        kpc::MagicItem mi;
        mi.design_slot = dslot;
        mi.target = tgt;
        mi.resid_slot = rs;
        mi.pad = 0;
        if (use_targets_) {
          kpc::CachedFit cf;
          if (kpc::target_cache().lookup(target_key(t->mask, tgt), &cf)) {
            mi.pad = 1 + (int)(w.fixed.size() / kpc::kMagicInfo);
            w.fixed.insert(w.fixed.end(), cf.record.begin(), cf.record.end());
            ++target_hits_;
          }
        }
        w.fit_items.push_back(mi);
        w.fit_mask.push_back(t->mask);
        w.comp_src.push_back((long long)rs * n_);
      }
      w.pair_x.push_back(slots[0]);
      w.pair_y.push_back(slots[1]);
      w.fit_x.push_back(rslot[0]);
      w.fit_y.push_back(rslot[1]);
      w.task_idx.push_back(ti);
    }
    if (uni) W[0].smooth_items.swap(shared_smooth);
    for (int li = 0; li < nlanes; ++li) {
      W[li].nsm = (int)W[li].smooth_items.size() / 4;
      stats_.fits += (long long)W[li].fit_items.size();
      stats_.components += (long long)W[li].comp_dst.size();
      stats_.setups += (long long)W[li].geo_items.size();
    }
    stats_.fits += (long long)claimed_fits_.size();
    stats_.setups += (long long)claimed_geos_.size();
    stats_.components += (long long)claimed_comps_.size();
    if (with_pairs) stats_.physical_tests += (long long)tasks.size();
    stage_host_ += now_s() - t0;
    const double* comp_source = level_ == 0 ? d_data_ : resid_.ptr;
    // uploads: a lane's work lists are packed into its page-locked block and
    // copied from there on the lane's stream, ahead of its launches (the
    // previous wave's copies from the block ended at its synchronisation)
    for (int li = 0; li < nlanes; ++li) {
      const Lane L = lane(li);
      const LaneWork& w = W[li];
      PinnedBuf& pb = res_->up[li];
      size_t need = 0;
      auto room = [&need](const auto& v) {
        need += (v.size() * sizeof(v[0]) + 255) & ~(size_t)255;
      };
      room(w.smooth_items);
      room(w.geo_items);
      room(w.fit_items);
      room(w.fixed);
      room(w.comp_src);
      room(w.comp_dst);
      room(w.pair_x);
      room(w.pair_y);
      if (li == 0) room(seeds);
      pb.reserve(need);
      size_t off = 0;
      auto put = [&](auto& dev, const auto& v) {
        dev.reserve(v.size());
        if (v.empty()) return;
        const size_t bytes = v.size() * sizeof(v[0]);
        std::memcpy(pb.ptr + off, v.data(), bytes);
        hip_check(hipMemcpyAsync(dev.ptr, pb.ptr + off, bytes, hipMemcpyHostToDevice, L.stream),
                  "hipMemcpyAsync H2D");
        off += (bytes + 255) & ~(size_t)255;
      };
      put(*L.smooth_items, w.smooth_items);
      put(*L.geo_items, w.geo_items);
      put(*L.fit_items, w.fit_items);
      put(*L.fixed, w.fixed);
      put(*L.comp_src, w.comp_src);
      put(*L.comp_dst, w.comp_dst);
      put(*L.pair_x, w.pair_x);
      put(*L.pair_y, w.pair_y);
      if (li == 0) put(d_seeds_, seeds);
      L.pair_out->reserve(w.pair_x.size() * kpc::kPairOut);
    }
    // designs on smooths built ahead on lane 1: lane 0 follows that launch
    if (wave_claimed_ && prefetch_pending_) {
      hip_check(hipStreamWaitEvent(stream_, res_->evp, 0), "stream wait");
      prefetch_pending_ = false;
    }
    // arenas built ahead on the third stream: both lanes follow that chain
    if (ahead_wait_) {
      hip_check(hipStreamWaitEvent(stream_, res_->evs[3], 0), "stream wait");
      hip_check(hipStreamWaitEvent(stream2_, res_->evs[3], 0), "stream wait");
      ahead_wait_ = false;
    }
    for (int li = 0; li < nlanes; ++li) {
      const Lane L = lane(li);
      LaneWork& w = W[li];
      hip_check(hipEventRecord(L.ev[0], L.stream), "event");
      hipGraph_t wave_graph = nullptr;
      if (graph)
        hip_check(hipStreamBeginCapture(L.stream, hipStreamCaptureModeRelaxed), "begin capture");
      if (w.nsm > 0) {
        const long long stride = kpc::smooth_scratch_doubles(n_);
        int grid = std::min(w.nsm, grid_limit(stride, 512));
        L.scratch->reserve((size_t)grid * stride);
        hip_check(kpc::launch_smooth_setup(d_data_, n_, p_, L.smooth_items->ptr, w.nsm,
                                           uni ? usmooth_X_.ptr : lsmooth_X_.ptr,
                                           uni ? usmooth_S_.ptr : lsmooth_S_.ptr,
                                           uni ? usmooth_info_.ptr : lsmooth_info_.ptr,
                                           L.scratch->ptr, grid, L.stream),
                  "smooth setup launch");
        pending_smooth_checks_ = true;
      }
      if (uni && nlanes > 1) {
        if (li == 0) hip_check(hipEventRecord(ev_[6], stream_), "event");
        else hip_check(hipStreamWaitEvent(L.stream, ev_[6], 0), "stream wait");
      }
      if (!graph) hip_check(hipEventRecord(L.ev[1], L.stream), "event");
      if (!w.geo_items.empty()) {
        const long long stride = kpc::geometry_scratch_doubles(q_);
        int grid = std::min((int)w.geo_items.size(), grid_limit(stride, 2048));
        L.scratch2->reserve((size_t)grid * stride);
        hip_check(kpc::launch_geometry(L.geo_items->ptr, (int)w.geo_items.size(), n_, q_, m_, dim_,
                                       rk_, uni ? usmooth_X_.ptr : lsmooth_X_.ptr,
                                       uni ? usmooth_S_.ptr : lsmooth_S_.ptr, design_.ptr,
                                       design_status_.ptr, (long long)n_ * 29, L.scratch2->ptr,
                                       grid, L.stream),
                  "geometry launch");
      }
      if (!graph) hip_check(hipEventRecord(L.ev[2], L.stream), "event");
      if (!w.fit_items.empty()) {
        const long long stride = kpc::magic_scratch_doubles(q_, m_, rk_, n_);
        int grid = std::min((int)w.fit_items.size(), grid_limit(stride, 65535));
        L.scratch3->reserve((size_t)grid * stride);
        hip_check(kpc::launch_magic(L.fit_items->ptr, (int)w.fit_items.size(), n_, q_, m_, rk_,
                                    d_data_, design_.ptr, resid_.ptr, rinfo_.ptr,
                                    L.scratch3->ptr, grid, ctl_, L.stream,
                                    w.fixed.empty() ? nullptr : L.fixed->ptr),
                  "magic launch");
      }
      if (!graph) hip_check(hipEventRecord(L.ev[3], L.stream), "event");
      if (!w.comp_dst.empty()) launch_components(L, comp_source, w.comp_dst.size());
      if (!graph) hip_check(hipEventRecord(L.ev[4], L.stream), "event");
      if (with_pairs) launch_pairs(L, (int)w.pair_x.size());
      if (graph) {
        hip_check(hipStreamEndCapture(L.stream, &wave_graph), "end capture");
        hipGraphExec_t exec = nullptr;
        hip_check(hipGraphInstantiate(&exec, wave_graph, nullptr, nullptr, 0), "graph instantiate");
        hip_check(hipGraphLaunch(exec, L.stream), "graph launch");
        hip_check(hipEventRecord(L.ev[5], L.stream), "event");
        hip_check(hipStreamSynchronize(L.stream), "graph synchronize");
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(wave_graph);
      } else {
        hip_check(hipEventRecord(L.ev[5], L.stream), "event");
      }
    }
    // the next level's work, queued behind this wave's launches
    if (ahead_queued_) launch_ahead();
    // read-backs into the page-locked block, queued behind the kernels: each
    // lane's test outputs on its own stream, then on lane 0 (after lane 1's
    // last kernel) the level's status words; one synchronisation of the lanes
    const std::vector<kpc::MagicItem>* fits = &W[0].fit_items;
    const std::vector<kpc::GeometryItem>* geos = &W[0].geo_items;
    const std::vector<int>* dst = &W[0].comp_dst;
    std::vector<kpc::MagicItem> fits2;
    std::vector<kpc::GeometryItem> geos2;
    std::vector<int> dst2;
    if (nlanes > 1 || !claimed_fits_.empty() || !claimed_geos_.empty()) {
      fits2 = W[0].fit_items;
      geos2 = W[0].geo_items;
      if (nlanes > 1) {
        fits2.insert(fits2.end(), W[1].fit_items.begin(), W[1].fit_items.end());
        geos2.insert(geos2.end(), W[1].geo_items.begin(), W[1].geo_items.end());
      }
      if (nlanes > 1 || !claimed_comps_.empty()) {
        dst2 = W[0].comp_dst;
        if (nlanes > 1) dst2.insert(dst2.end(), W[1].comp_dst.begin(), W[1].comp_dst.end());
        dst2.insert(dst2.end(), claimed_comps_.begin(), claimed_comps_.end());
        dst = &dst2;
      }
      fits2.insert(fits2.end(), claimed_fits_.begin(), claimed_fits_.end());
      geos2.insert(geos2.end(), claimed_geos_.begin(), claimed_geos_.end());
      fits = &fits2;
      geos = &geos2;
    }
    const bool hsic = method_ == KPC_HSIC_GAMMA || method_ == KPC_HSIC_PERM;
    const StatusReadBack rb = plan_status(*fits, *geos, W[0].nsm + W[1].nsm + (wave_claimed_ ? 1 : 0), dst, true,
                                          with_pairs && hsic, 0);
    size_t out_off[2] = {0, 0};
    size_t down_end = rb.end;
    for (int li = 0; li < nlanes && with_pairs; ++li) {
      out_off[li] = down_end;
      down_end += aligned256(W[li].pair_x.size() * kpc::kPairOut * sizeof(double));
    }
    res_->down.reserve(down_end);
    for (int li = 0; li < nlanes && with_pairs; ++li) {
      const Lane L = lane(li);
      const size_t bytes = W[li].pair_x.size() * kpc::kPairOut * sizeof(double);
      if (bytes)
        hip_check(hipMemcpyAsync(res_->down.ptr + out_off[li], L.pair_out->ptr, bytes,
                                 hipMemcpyDeviceToHost, L.stream),
                  "D2H p-values");
    }
    if (nlanes > 1) hip_check(hipStreamWaitEvent(stream_, ev2_[5], 0), "stream wait");
    queue_status(rb, stream_);
    for (int li = nlanes - 1; li >= 0; --li)
      hip_check(hipStreamSynchronize(lane(li).stream), "wave synchronize");
    // stage times are summed over the lanes (they overlap in wall time)
    float ms;
    account_ahead();
    account_prefetch();
    for (int li = 0; li < nlanes; ++li) {
      const Lane L = lane(li);
      const LaneWork& w = W[li];
      const bool launched[5] = {w.nsm > 0, !w.geo_items.empty(), !w.fit_items.empty(),
                                !w.comp_dst.empty(), with_pairs};
      if (graph) {
        // one graph per wave: only the whole pipeline is timed
        hip_check(hipEventElapsedTime(&ms, L.ev[0], L.ev[5]), "event elapsed");
        graph_seconds_ += ms * 1e-3;
        ++graph_launches_;
      }
      for (int s = 0; s < 5 && !graph; ++s) {
        hip_check(hipEventElapsedTime(&ms, L.ev[s], L.ev[s + 1]), "event elapsed");
        stage_[s] += ms * 1e-3;
      }
      for (int s = 0; s < 5; ++s)
        if (launched[s]) ++launches_[s];
    }
    const double* rinfo_h = verify_status(rb, *fits, *geos, dst);
    if (use_targets_ && rinfo_h)
      for (int li = 0; li < nlanes; ++li)
        for (size_t i = 0; i < W[li].fit_items.size(); ++i) {
          const kpc::MagicItem& f = W[li].fit_items[i];
          if (f.pad > 0) continue;
          kpc::CachedFit cf;
          std::memcpy(cf.record.data(), rinfo_h + (size_t)f.resid_slot * kpc::kMagicInfo,
                      sizeof(double) * kpc::kMagicInfo);
          if (m_ <= kpc::kMagicInfo - 10)  // the record holds every log sp
            kpc::target_cache().insert(target_key(W[li].fit_mask[i], f.target), cf);
        }
    if (!with_pairs) return;
    const double* cmargin_h =
        rb.cmargin_n ? reinterpret_cast<const double*>(res_->down.ptr + rb.cmargin_off) : nullptr;
    for (int li = 0; li < nlanes; ++li) {
      const LaneWork& w = W[li];
      const double* out = reinterpret_cast<const double*>(res_->down.ptr + out_off[li]);
      for (size_t j = 0; j < w.task_idx.size(); ++j) {
        Task* t = tasks[w.task_idx[j]];
        const double* o = out + kpc::kPairOut * j;
        TestResult res{o[0], o[1], o[2], o[3], o[4]};
        if (std::isnan(res.p)) res.p = 1.0;  // NAdelete = TRUE
        for (int rs : {w.fit_x[j], w.fit_y[j]})
          if (rs >= 0)
            res.margin = std::min(res.margin,
                                  rinfo_h[(size_t)rs * kpc::kMagicInfo + kpc::kMagicInfo - 2]);
        if (cmargin_h)
          res.margin = std::min(res.margin, std::min(cmargin_h[w.pair_x[j]], cmargin_h[w.pair_y[j]]));
        apply_result(*t, res);
        if (!rmode_) twin_map_.emplace(twin_key(*t), res);
        if (use_results_) {
          // both orientations: outside the R stream a test's value does not
          // depend on which variable conditions first
          const kpc::CachedResult cr{res.p, res.stat, res.mean, res.var, res.margin};
          Task twin = *t;
          std::swap(twin.x, twin.y);
          kpc::result_cache().insert(result_key(*t), cr);
          kpc::result_cache().insert(result_key(twin), cr);
        }
      }
    }
    for (auto& d : dups) apply_result(*d.first, *twin_map_.find(twin_key(*tasks[d.second])));
  }

  kpc::CacheKey result_key(const Task& t) const {
    kpc::CacheKey k;
    k.dataset = dataset_;
    k.a = opt_hash_;
    k.b = ((uint64_t)(uint32_t)t.x << 32) | (uint32_t)t.y;
    k.c = t.mask;
    return k;
  }
  kpc::CacheKey target_key(uint64_t mask, int target) const {
    kpc::CacheKey k;
    k.dataset = dataset_;
    k.a = 0x746172676574ULL;  // a fit, not a test
    k.b = mask;
    k.c = (uint64_t)target;
    return k;
  }
  long long result_hits() const { return result_hits_; }
  long long target_hits() const { return target_hits_; }

  static std::pair<uint64_t, uint64_t> twin_key(const Task& t) {
    const uint64_t a = (uint64_t)std::min(t.x, t.y), b = (uint64_t)std::max(t.x, t.y);
    return std::make_pair(t.mask, (a << 32) | b);
  }

  static void apply_result(Task& t, const TestResult& r) {
    t.p = r.p;
    t.stat = r.stat;
    t.mean = r.mean;
    t.var = r.var;
    t.margin = r.margin;
    t.computed = true;
  }

  // One launch lane: a stream, its six stage events, and the work lists and
  // scratch slabs its kernels use.  Lane 0 is the engine's own stream; tests
  // split by conditioning set run on lanes 0 and 1 at once (their designs,
  // fits, components and tests are disjoint), so the last setups or the
  // slowest fit of one lane overlap the other lane's later stages.
  struct Lane {
    hipStream_t stream;
    hipEvent_t* ev;
    DevArray<int>* smooth_items;
    DevArray<kpc::GeometryItem>* geo_items;
    DevArray<kpc::MagicItem>* fit_items;
    DevArray<long long>* comp_src;
    DevArray<int>* comp_dst;
    DevArray<int>* pair_x;
    DevArray<int>* pair_y;
    DevArray<double>* pair_out;
    DevArray<double>* scratch;   // smooth setups
    DevArray<double>* scratch2;  // geometry
    DevArray<double>* scratch3;  // GCV fits
    DevArray<double>* scratch4;  // components
    DevArray<double>* fixed;     // cached fit records
  };
  Lane lane(int i) {
    if (i == 0)
      return Lane{stream_,        ev_,          &d_smooth_items_, &d_geo_items_, &d_fit_items_,
                  &d_comp_src_,   &d_comp_dst_, &d_pair_x_,       &d_pair_y_,    &d_pair_out_,
                  &scratch_,      &scratch2_,   &scratch3_,       &scratch4_,
                  &d_fixed_};
    return Lane{stream2_,        ev2_,          &l1_smooth_items_, &l1_geo_items_, &l1_fit_items_,
                &l1_comp_src_,   &l1_comp_dst_, &l1_pair_x_,       &l1_pair_y_,    &l1_pair_out_,
                &l1_scratch_,    &l1_scratch2_, &l1_scratch3_,     &l1_scratch4_,
                &l1_fixed_};
  }

  void launch_pairs(int count) { launch_pairs(lane(0), count); }
  void launch_pairs(const Lane& L, int count) {
    const int grid = std::min(count, 65535);
    switch (method_) {
      case KPC_DCC_GAMMA:
        hip_check(kpc::launch_dcov_pairs(L.pair_x->ptr, L.pair_y->ptr, count, n_, k_,
                                         comp_vals_.ptr, comp_vecs_.ptr, comp_dsum_.ptr,
                                         comp_mom_.ptr, L.pair_out->ptr, grid, L.stream),
                  "dcov pair launch");
        break;
      case KPC_HSIC_GAMMA:
        hip_check(kpc::launch_hsic_pairs(L.pair_x->ptr, L.pair_y->ptr, count, n_, k_,
                                         comp_vecs_.ptr, comp_status_.ptr, comp_dsum_.ptr,
                                         comp_mom_.ptr, L.pair_out->ptr, grid, L.stream),
                  "hsic pair launch");
        break;
      default:
        perm_pairs(count);
        break;
    }
  }

  int hsic_kt() const { return (k_ + 15) / 16; }

  // Permutation tests of `count` pairs: observed statistics (identity
  // permutation), then chunks of at most perm_chunk_replicates() replicates:
  // generate the index arrays, evaluate every replicate, tally per test in
  // replicate order; finally p, statistic, mean, variance and margin.
  long long perm_chunk_replicates() const {
    const long long b = kPermScratchBytes / (2LL * n_);
    return b < 1 ? 1 : b;
  }

  void perm_values(const kpc::PermUnit* units, int nunits, const unsigned short* perms,
                   double* vals, int block) {
    const int grid = std::min(nunits, 65535);
    if (method_ == KPC_HSIC_PERM)
      hip_check(kpc::launch_hsic_perm_values(units, nunits, d_pair_x_.ptr, d_pair_y_.ptr, n_,
                                             hsic_kt(), comp_rows_.ptr, perms, vals, grid,
                                             stream_),
                "hsic.perm replicate launch");
    else
      hip_check(kpc::launch_dcc_perm_values(units, nunits, d_pair_x_.ptr, d_pair_y_.ptr, n_,
                                            n_ * comp_k_, index_, comp_vecs_.ptr, comp_dsum_.ptr,
                                            perms, vals, dcc_perm_jt_, perms ? block : 1,
                                            grid, stream_),
                "dcc.perm replicate launch");
  }

  void perm_pairs(int count) {
    const bool hsic = method_ == KPC_HSIC_PERM;
    if (rmode_)
      for (int t = 0; t < count; ++t)
        if ((size_t)t >= pair_rtab_.size() || pair_rtab_[t] < 0)
          throw KpcError(-4, "R-stream test has no drawn permutation table");
    const int R = opt_.replicates;
    const double scale = hsic ? 1.0 : 1.0 / (double)n_;
    const long long budget = perm_chunk_replicates();
    std::vector<kpc::PermUnit> obs_units((size_t)count);
    std::vector<double> state((size_t)4 * count);
    for (int t = 0; t < count; ++t) {
      obs_units[t] = {t, 0, 1, t};
      state[4 * t] = opt_.include_observed ? 1.0 : 0.0;
      state[4 * t + 1] = 0.0;
      state[4 * t + 2] = 0.0;
      state[4 * t + 3] = 1e300;
    }
    d_perm_units_.upload(obs_units, stream_);
    perm_state_.upload(state, stream_);
    perm_obs_.reserve((size_t)count);
    perm_values(d_perm_units_.ptr, count, nullptr, perm_obs_.ptr, 1);
    hip_check(hipStreamSynchronize(stream_), "permutation synchronize");
    if (R > 0) {
      const long long chunk_reps = std::min(budget, (long long)count * R);
      perm_scratch_.reserve((size_t)chunk_reps * n_);
      perm_vals_.reserve((size_t)chunk_reps);
      // tuned launch shapes (bit-identical results for every candidate):
      // replicates per generation unit and per replicate-kernel unit
      kpc::Tuner& tuner = kpc::Tuner::instance();
      const int nb = (n_ + 255) / 256 * 256;
      const int gmax = kpc::perm_generate_width(n_);
      const std::string gkey = "perm_generate_width/n" + std::to_string(nb);
      std::vector<int> gcand = gmax > 1 ? std::vector<int>{gmax, (gmax + 1) / 2}
                                        : std::vector<int>{1};
      if (const char* e = std::getenv("KPC_PERM_GEN_WIDTH")) {
        const int w = std::atoi(e);
        if (w >= 1 && w <= gmax) gcand = {w};
      }
      std::string vkey;
      std::vector<int> vcand;
      if (hsic) {
        vkey = "hsic_perm_block/n" + std::to_string(nb) + "/kt" + std::to_string(hsic_kt());
        vcand = {16, 8, 32};
        if (const char* e = std::getenv("KPC_HSIC_PERM_BLOCK")) {
          const int b = std::atoi(e);
          if (b >= 1 && b <= 256) vcand = {b};
        }
      } else {
        vkey = "dcc_perm_rb/n" + std::to_string(nb) + "/jt" + std::to_string(dcc_perm_jt_);
        vcand = dcc_shape_forced_ ? std::vector<int>{dcc_perm_rb_} : std::vector<int>{8, 12};
      }
      int gw = 0, vb = 0;
      bool gtrial = false, vtrial = false;
      auto pick = [&]() {
        gw = tuner.choose(gkey, gcand, &gtrial);
        vb = tuner.choose(vkey, vcand, &vtrial);
      };
      pick();
      // one chunk: [segments | generation units | value units]; the host list
      // is the source of an asynchronous copy, so each chunk ends with a
      // stream synchronisation before the list is rebuilt
      std::vector<kpc::PermUnit> segs, gen, val;
      long long used = 0;
      auto flush = [&]() {
        if (segs.empty()) return;
        const size_t ns = segs.size(), ng = gen.size(), nv = val.size();
        segs.insert(segs.end(), gen.begin(), gen.end());
        segs.insert(segs.end(), val.begin(), val.end());
        d_perm_units_.upload(segs, stream_);
        const kpc::PermUnit* U = d_perm_units_.ptr;
        hip_check(hipEventRecord(tev_[0], stream_), "event");
        std::vector<unsigned short> staged;
        if (rmode_) {
          // the R stream's tables, drawn on the host in execution order:
          // replicates r .. r + take of test t go to slots slot0 ..
          staged.resize((size_t)used * n_);
          for (size_t i = 0; i < ns; ++i) {
            const kpc::PermUnit& sg = segs[i];
            std::memcpy(staged.data() + (size_t)sg.slot * n_,
                        rtab_host_.data() + pair_rtab_[sg.test] + (size_t)sg.r0 * n_,
                        sizeof(unsigned short) * (size_t)sg.nr * n_);
          }
          hip_check(hipMemcpyAsync(perm_scratch_.ptr, staged.data(),
                                   sizeof(unsigned short) * staged.size(), hipMemcpyHostToDevice,
                                   stream_),
                    "R-stream permutation upload");
        } else {
          hip_check(kpc::launch_perm_generate(U + ns, (int)ng, d_seeds_.ptr, n_, gw,
                                              perm_scratch_.ptr, (int)std::min<size_t>(ng, 8192),
                                              stream_),
                    "permutation generate launch");
        }
        hip_check(hipEventRecord(tev_[1], stream_), "event");
        perm_values(U + ns + ng, (int)nv, perm_scratch_.ptr, perm_vals_.ptr, vb);
        hip_check(hipEventRecord(tev_[2], stream_), "event");
        hip_check(kpc::launch_perm_tally(U, (int)ns, perm_vals_.ptr, perm_obs_.ptr, scale,
                                         perm_state_.ptr, stream_),
                  "permutation tally launch");
        hip_check(hipStreamSynchronize(stream_), "permutation synchronize");
        // a trial counts when the chunk carried enough work to time
        const double reps = (double)used;
        if ((gtrial || vtrial) && reps * n_ >= 4e6) {
          float ms = 0.f;
          if (gtrial && !rmode_) {
            hip_check(hipEventElapsedTime(&ms, tev_[0], tev_[1]), "event elapsed");
            tuner.record(gkey, gw, ms * 1e-3, reps);
          }
          if (vtrial) {
            hip_check(hipEventElapsedTime(&ms, tev_[1], tev_[2]), "event elapsed");
            tuner.record(vkey, vb, ms * 1e-3, reps);
          }
        }
        segs.clear();
        gen.clear();
        val.clear();
        used = 0;
        pick();
      };
      for (int t = 0; t < count; ++t) {
        int r = 0;
        while (r < R) {
          const int take = (int)std::min<long long>(R - r, budget - used);
          const int slot0 = (int)used;
          segs.push_back({t, r, take, slot0});
          for (int b = 0; b < take; b += gw)
            gen.push_back({t, r + b, std::min(gw, take - b), slot0 + b});
          for (int b = 0; b < take; b += vb)
            val.push_back({t, r + b, std::min(vb, take - b), slot0 + b});
          used += take;
          r += take;
          if (used == budget) flush();
        }
      }
      flush();
    }
    hip_check(kpc::launch_perm_finalize(count, perm_obs_.ptr, perm_state_.ptr, scale, R,
                                        opt_.include_observed, hsic ? 0 : 1, d_pair_out_.ptr,
                                        stream_),
              "permutation finalize launch");
  }

  void launch_components(const Lane& L, const double* src, size_t count) {
    const long long* csrc = L.comp_src->ptr;
    const int* cdst = L.comp_dst->ptr;
    if (method_ == KPC_HSIC_GAMMA || method_ == KPC_HSIC_PERM) {
      const int grid = std::min((int)count, 8192);
      hip_check(kpc::launch_inchol_components(src, csrc, cdst, (int)count, n_, k_,
                                              1.0 / opt_.hsic_sig, comp_vecs_.ptr,
                                              comp_dsum_.ptr, comp_mom_.ptr, comp_status_.ptr,
                                              comp_margin_.ptr,
                                              method_ == KPC_HSIC_PERM ? comp_rows_.ptr : nullptr,
                                              16 * hsic_kt(), grid, L.stream),
                "inchol launch");
      return;
    }
    if (method_ == KPC_DCC_PERM) {
      const int grid = std::min((int)count, 8192);
      hip_check(kpc::launch_dcc_rowmeans(src, csrc, cdst, (int)count, n_, n_ * comp_k_, index_,
                                         comp_vecs_.ptr, comp_dsum_.ptr, grid, L.stream),
                "dcc row-mean launch");
      return;
    }
    const int mmax = kpc::lanczos_mmax(n_);
    if (dense_components_) {
      const long long stride = kpc::dcov_component_scratch_doubles(n_, k_);
      int grid = std::min((int)count, grid_limit(stride, 2048));
      L.scratch4->reserve((size_t)grid * stride);
      hip_check(kpc::launch_dcov_components(src, csrc, cdst, (int)count, n_, k_, comp_vals_.ptr,
                                            comp_vecs_.ptr, comp_dsum_.ptr, comp_mom_.ptr,
                                            L.scratch4->ptr, grid, L.stream),
                "dcov component launch");
    } else {
      const long long stride = kpc::dcov_lanczos_scratch_doubles(n_, k_, mmax);
      int grid = std::min((int)count, grid_limit(stride, 4096));
      L.scratch4->reserve((size_t)grid * stride);
      hip_check(kpc::launch_dcov_lanczos(src, csrc, cdst, (int)count, n_, k_, comp_vals_.ptr,
                                         comp_vecs_.ptr, comp_dsum_.ptr, comp_mom_.ptr,
                                         comp_status_.ptr, L.scratch4->ptr, grid, mmax, 1e-13,
                                         L.stream),
                "dcov lanczos launch");
    }
  }

  int level() const { return level_; }
  int n() const { return n_; }
  int p() const { return p_; }
  double stage(int i) const { return stage_[i]; }
  long long launches(int i) const { return launches_[i]; }
  double stage_host() const { return stage_host_; }
  double graph_seconds() const { return graph_seconds_; }
  long long graph_launches() const { return graph_launches_; }
  struct Stats {
    long long physical_tests = 0, fits = 0, components = 0, setups = 0, evals = 0,
              lanczos_steps = 0;
  } stats_;

  // direct access for the regrXonS / gcv entry points
  double* data_ptr() { return d_data_; }
  hipStream_t stream() { return stream_; }
  const kpc::MagicControl& ctl() const { return ctl_; }

  // Build designs for arbitrary S sets (all |S| == level_); returns slots.
  std::vector<int> build_designs(const std::vector<Task>& sets) {
    std::vector<int> smooth_items;
    std::vector<kpc::GeometryItem> geo_items;
    std::vector<int> slots;
    for (const Task& t : sets) slots.push_back(ensure_design(t, smooth_items, geo_items));
    d_smooth_items_.upload(smooth_items, stream_);
    d_geo_items_.upload(geo_items, stream_);
    const int nsm = (int)smooth_items.size() / 4;
    const bool uni = uni_smooths();
    if (nsm > 0) {
      const long long stride = kpc::smooth_scratch_doubles(n_);
      int grid = std::min(nsm, grid_limit(stride, 512));
      scratch_.reserve((size_t)grid * stride);
      hip_check(kpc::launch_smooth_setup(d_data_, n_, p_, d_smooth_items_.ptr, nsm,
                                         uni ? usmooth_X_.ptr : lsmooth_X_.ptr,
                                         uni ? usmooth_S_.ptr : lsmooth_S_.ptr,
                                         uni ? usmooth_info_.ptr : lsmooth_info_.ptr,
                                         scratch_.ptr, grid, stream_),
                "smooth setup launch");
    }
    if (!geo_items.empty()) {
      const long long stride = kpc::geometry_scratch_doubles(q_);
      int grid = std::min((int)geo_items.size(), grid_limit(stride, 2048));
      scratch2_.reserve((size_t)grid * stride);
      hip_check(kpc::launch_geometry(d_geo_items_.ptr, (int)geo_items.size(), n_, q_, m_, dim_, rk_,
                                     uni ? usmooth_X_.ptr : lsmooth_X_.ptr,
                                     uni ? usmooth_S_.ptr : lsmooth_S_.ptr, design_.ptr,
                                     design_status_.ptr, (long long)n_ * 29, scratch2_.ptr, grid,
                                     stream_),
                "geometry launch");
    }
    hip_check(hipStreamSynchronize(stream_), "setup synchronize");
    check_device_status({}, geo_items, nsm);
    return slots;
  }
  double* design_ptr() { return design_.ptr; }
  int slab_grid(long long stride, int cap) const { return grid_limit(stride, cap); }
  int q() const { return q_; }
  int m() const { return m_; }
  int rk() const { return rk_; }
  DevArray<double>& resid_arena() { return resid_; }
  DevArray<double>& rinfo_arena() { return rinfo_; }
  DevArray<double>& scratch3() { return scratch3_; }

 private:
  int grid_limit(long long stride, int cap) const {
    // keep any one scratch slab under ~8 GiB
    long long lim = (8LL << 30) / 8 / std::max(1LL, stride);
    return (int)std::max(1LL, std::min((long long)cap, lim));
  }

  int ensure_design(const Task& t, std::vector<int>& smooth_items,
                    std::vector<kpc::GeometryItem>& geo_items) {
    if (const int* it = design_map_.find(t.mask)) return *it;
    if (ahead_claim_) {
      if (const int* ahead = claim_design_map_.find(t.mask)) {
        // built ahead during the previous level on these columns' smooths
        const int ds = *ahead;
        design_map_.set(t.mask, ds);
        kpc::GeometryItem gi;
        std::memset(&gi, 0, sizeof(gi));
        gi.design_slot = ds;
        for (int i = 0; i < t.slen; ++i) {
          const int col = t.S[i];
          if (uni_slot_[col] < 0) {
            uni_slot_[col] = col;
            if (uni_ready_[col]) {
              uni_ready_[col] = 0;  // built ahead on lane 1
              wave_claimed_ = true;
            }
          }
          gi.smooth_slot[i] = col;
        }
        claimed_geos_.push_back(gi);
        return ds;
      }
    }
    if (l1_claim_) {
      // built ahead: design slot s, on column s's univariate smooth
      const int c = t.S[0];
      design_map_.set(t.mask, c);
      if (uni_slot_[c] < 0) {
        uni_slot_[c] = c;
        uni_ready_[c] = 0;
      }
      wave_claimed_ = true;
      kpc::GeometryItem gi;
      std::memset(&gi, 0, sizeof(gi));
      gi.design_slot = c;
      gi.smooth_slot[0] = c;
      claimed_geos_.push_back(gi);
      return c;
    }
    const int ds = n_design_++;
    design_map_.set(t.mask, ds);
    kpc::GeometryItem gi;
    std::memset(&gi, 0, sizeof(gi));
    gi.design_slot = ds;
    if (!uni_smooths()) {
      const int ss = n_lsmooth_++;
      smooth_items.push_back(level_);
      smooth_items.push_back(t.S[0]);
      smooth_items.push_back(level_ == 2 ? t.S[1] : t.S[0]);
      smooth_items.push_back(ss);
      gi.smooth_slot[0] = ss;
    } else {
      for (int i = 0; i < t.slen; ++i) {
        const int col = t.S[i];
        if (uni_slot_.empty()) uni_slot_.assign(p_, -1);
        if (uni_slot_[col] < 0) {
          uni_slot_[col] = col;
          if (!uni_ready_.empty() && uni_ready_[col]) {
            uni_ready_[col] = 0;  // built ahead
            wave_claimed_ = true;
          } else {
            smooth_items.push_back(1);
            smooth_items.push_back(col);
            smooth_items.push_back(col);
            smooth_items.push_back(col);
          }
        }
        gi.smooth_slot[i] = col;
      }
    }
    geo_items.push_back(gi);
    return ds;
  }

  // Status read-back of a wave: smooth set-up info, the design status words
  // (one slot range), the fit records of every residual slot of the level,
  // component status and HSIC component margins.  plan_status lays the
  // regions out in the page-locked read-back block from byte `base` on,
  // queue_status copies them there on one stream (behind that stream's
  // kernels), and verify_status checks them once that stream is synchronised.
  struct StatusReadBack {
    size_t end = 0;                          // first byte past the regions
    size_t smooth_off = 0, smooth_n = 0;     // ints, 4 per smooth slot
    size_t design_off = 0, design_n = 0;     // doubles
    int design_lo = 0;
    size_t rinfo_off = 0, rinfo_n = 0;       // doubles, kMagicInfo per residual slot
    size_t cstat_off = 0, cstat_n = 0;       // ints
    size_t cmargin_off = 0, cmargin_n = 0;   // doubles
    bool smooth_uni = false;
  };
  static size_t aligned256(size_t b) { return (b + 255) & ~(size_t)255; }

  StatusReadBack plan_status(const std::vector<kpc::MagicItem>& fits,
                             const std::vector<kpc::GeometryItem>& geos, int nsm,
                             const std::vector<int>* dst, bool want_info, bool want_margins,
                             size_t base) const {
    StatusReadBack rb;
    size_t off = base;
    if (nsm > 0) {
      rb.smooth_uni = uni_smooths();
      rb.smooth_n = (rb.smooth_uni ? (size_t)p_ : (size_t)n_lsmooth_) * 4;
      rb.smooth_off = off;
      off += aligned256(rb.smooth_n * sizeof(int));
    }
    if (!geos.empty()) {
      int lo = geos[0].design_slot, hi = lo;
      for (const auto& g : geos) {
        lo = std::min(lo, g.design_slot);
        hi = std::max(hi, g.design_slot);
      }
      rb.design_lo = lo;
      rb.design_n = (size_t)(hi - lo + 1);
      rb.design_off = off;
      off += aligned256(rb.design_n * sizeof(double));
    }
    if (!fits.empty() || (want_info && level_ > 0 && n_resid_ > 0)) {
      rb.rinfo_n = (size_t)n_resid_ * kpc::kMagicInfo;
      rb.rinfo_off = off;
      off += aligned256(rb.rinfo_n * sizeof(double));
    }
    if (dst && !dst->empty() && method_ != KPC_DCC_PERM &&
        !(method_ == KPC_DCC_GAMMA && dense_components_)) {
      rb.cstat_n = (size_t)n_comp_;
      rb.cstat_off = off;
      off += aligned256(rb.cstat_n * sizeof(int));
    }
    if (want_margins) {
      rb.cmargin_n = (size_t)n_comp_;
      rb.cmargin_off = off;
      off += aligned256(rb.cmargin_n * sizeof(double));
    }
    rb.end = off;
    return rb;
  }

  // the read-back block must already hold rb.end bytes
  void queue_status(const StatusReadBack& rb, hipStream_t s) {
    unsigned char* h = res_->down.ptr;
    auto down = [&](size_t off, const void* src, size_t bytes) {
      if (bytes)
        hip_check(hipMemcpyAsync(h + off, src, bytes, hipMemcpyDeviceToHost, s), "status read-back");
    };
    down(rb.smooth_off, (rb.smooth_uni ? usmooth_info_ : lsmooth_info_).ptr,
         rb.smooth_n * sizeof(int));
    down(rb.design_off, design_status_.ptr + rb.design_lo, rb.design_n * sizeof(double));
    down(rb.rinfo_off, rinfo_.ptr, rb.rinfo_n * sizeof(double));
    down(rb.cstat_off, comp_status_.ptr, rb.cstat_n * sizeof(int));
    down(rb.cmargin_off, comp_margin_.ptr, rb.cmargin_n * sizeof(double));
  }

  // Returns the fit records of the level's residual slots (in the read-back
  // block, valid until the next read-back) or null when none were read.
  const double* verify_status(const StatusReadBack& rb, const std::vector<kpc::MagicItem>& fits,
                              const std::vector<kpc::GeometryItem>& geos,
                              const std::vector<int>* dst) {
    const unsigned char* h = res_->down.ptr;
    if (rb.smooth_n) {
      const int* info = reinterpret_cast<const int*>(h + rb.smooth_off);
      for (size_t i = 0; i < rb.smooth_n / 4; ++i) {
        if (rb.smooth_uni && uni_slot_[i] < 0) continue;
        if (info[4 * i] != 0)
          throw KpcError(-5, "thin-plate setup failed (fewer unique locations than basis size)");
      }
    }
    if (rb.design_n) {
      const double* st = reinterpret_cast<const double*>(h + rb.design_off);
      for (const auto& g : geos)
        if (st[(size_t)(g.design_slot - rb.design_lo)] != 0.0)
          throw KpcError(-5, "design geometry failed (penalty rank or initial.sp)");
    }
    const double* rinfo = rb.rinfo_n ? reinterpret_cast<const double*>(h + rb.rinfo_off) : nullptr;
    for (const auto& f : fits) {
      const double* o = rinfo + (size_t)f.resid_slot * kpc::kMagicInfo;
      if (o[0] != 0.0) throw KpcError(-6, "GCV optimiser did not converge");
      if (f.pad == 0) stats_.evals += (long long)o[5];
    }
    if (rb.cstat_n) {
      const int* cs = reinterpret_cast<const int*>(h + rb.cstat_off);
      for (int sl : *dst) {
        if (method_ != KPC_DCC_GAMMA) {
          if (cs[sl] <= 0 || cs[sl] > k_) throw KpcError(-7, "incomplete Cholesky produced no factor");
          continue;
        }
        if (cs[sl] <= 0) throw KpcError(-7, "dcov component eigensolver did not converge");
        stats_.lanczos_steps += cs[sl];
      }
    }
    return rinfo;
  }

  // Status words of set-ups run on the engine's own stream.
  void check_device_status(const std::vector<kpc::MagicItem>& fits,
                           const std::vector<kpc::GeometryItem>& geos, int nsm) {
    const StatusReadBack rb = plan_status(fits, geos, nsm, nullptr, false, false, 0);
    res_->down.reserve(rb.end);
    queue_status(rb, stream_);
    hip_check(hipStreamSynchronize(stream_), "status synchronize");
    verify_status(rb, fits, geos, nullptr);
  }

  int n_, p_, k_;
  int method_ = KPC_DCC_GAMMA;
  int comp_k_ = 0;  // doubles per row of a component slot (n * comp_k_ per slot)
  double index_ = 1.0;
  uint64_t seed_ = 0;
  kpc_options opt_;
  EngineHipResources* res_ = nullptr;
  hipStream_t stream_ = nullptr, stream2_ = nullptr;
  int lanes_ = 2;
  bool share_uni_l1_ = true;  // level-1 smooths in the per-column cache
  std::vector<char> uni_ready_;  // univariate smooths built ahead, not yet claimed
  bool prefetch_pending_ = false;  // lane 0 not yet ordered after that launch
  bool prefetch_time_pending_ = false;  // its stage times not yet counted
  bool wave_claimed_ = false;      // the current wave claimed one (status read-back)
  DevArray<int> pre_items_;
  DevArray<double> pre_scratch_;
  DevArray<kpc::GeometryItem> pre_geo_items_;
  DevArray<kpc::MagicItem> pre_fit_items_;
  DevArray<double> pre_scratch2_, pre_scratch3_;
  bool l1_ahead_ = false;  // level 1's designs and fits built ahead
  bool l1_claim_ = false;  // this level claims them (level 1, adopted)
  // the next level's designs and fits built ahead on the third stream
  hipStream_t stream3_ = nullptr;
  int ahead_level_ = -1;         // the level the lists are for
  bool ahead_queued_ = false;    // lists built, launch behind the next wave
  bool ahead_launched_ = false;  // a chain was launched (evs[2] recorded)
  bool ahead_ready_ = false;     // launched for ahead_level_, not yet adopted
  bool ahead_claim_ = false;     // this level claims them
  bool ahead_wait_ = false;      // lanes not yet ordered after the chain
  std::vector<kpc::GeometryItem> ahead_geo_;
  std::vector<kpc::MagicItem> ahead_fits_;
  bool ahead_time_pending_ = false;  // an adopted chain's stage times
  CompSet chain_comp_;  // components of the chain built ahead
  DevArray<long long> chain_comp_src_;
  DevArray<int> chain_comp_dst_;
  DevArray<double> chain_scratch4_;
  bool comps_ahead_ = true;    // chains build their fits' components
  bool chain_comps_ = false;   // the last chain launched did
  bool l1_comps_ = false;      // the level-1 chain did
  bool claim_comps_ = false;   // this level claims components built ahead
  std::vector<int> claimed_comps_;  // the wave's claimed components (status)
  FlatMap<uint64_t, int, U64Hash> ahead_design_map_, claim_design_map_;
  FlatMap<std::pair<int, uint64_t>, int, KeyHash> ahead_resid_map_, claim_resid_map_;
  DevArray<double> ahead_design_, ahead_design_status_, ahead_resid_, ahead_rinfo_;
  DevArray<kpc::GeometryItem> ahead_geo_items_;
  DevArray<kpc::MagicItem> ahead_fit_items_;
  DevArray<double> ahead_scratch2_, ahead_scratch3_;
  std::vector<kpc::GeometryItem> claimed_geos_;  // the wave's claimed designs
  std::vector<kpc::MagicItem> claimed_fits_;     // and fits (status read-back)
  double* d_data_ = nullptr;
  size_t data_bytes_ = 0;
  hipEvent_t ev_[8];
  kpc::MagicControl ctl_;
  int level_ = -1, q_ = 0, m_ = 0, dim_ = 0, rk_ = 0;
  kpc::DesignLayout layout_{};
  DevArray<double> design_status_;  // status word of each design slot (compact copy)
  FlatMap<uint64_t, int, U64Hash> design_map_;
  FlatMap<std::pair<int, uint64_t>, int, KeyHash> resid_map_, comp_map_;
  FlatMap<uint64_t, int, U64Hash> lane_map_;  // conditioning set -> lane (one wave)
  FlatMap<std::pair<uint64_t, uint64_t>, TestResult, TwinHash> twin_map_;
  FlatMap<std::pair<uint64_t, uint64_t>, size_t, TwinHash> in_batch_;
  std::vector<int> uni_slot_;
  int n_design_ = 0, n_resid_ = 0, n_comp_ = 0, n_lsmooth_ = 0;
  DevArray<double> lsmooth_X_, lsmooth_S_, usmooth_X_, usmooth_S_;
  DevArray<int> lsmooth_info_, usmooth_info_;
  DevArray<double> design_, resid_, rinfo_;
  DevArray<double> comp_vals_, comp_vecs_, comp_dsum_, comp_mom_;
  DevArray<int> comp_status_;
  DevArray<double> comp_margin_;
  DevArray<unsigned long long> d_seeds_;
  bool dense_components_ = std::getenv("KPC_DCOV_DENSE") != nullptr;
  DevArray<double> scratch_, scratch2_, scratch3_, scratch4_;
  DevArray<int> d_smooth_items_;
  DevArray<kpc::GeometryItem> d_geo_items_;
  DevArray<kpc::MagicItem> d_fit_items_;
  DevArray<long long> d_comp_src_;
  DevArray<int> d_comp_dst_, d_pair_x_, d_pair_y_;
  DevArray<double> d_pair_out_;
  // the second launch lane's work lists and scratch slabs
  hipEvent_t ev2_[6];
  DevArray<int> l1_smooth_items_;
  DevArray<kpc::GeometryItem> l1_geo_items_;
  DevArray<kpc::MagicItem> l1_fit_items_;
  DevArray<long long> l1_comp_src_;
  DevArray<int> l1_comp_dst_, l1_pair_x_, l1_pair_y_;
  DevArray<double> l1_pair_out_;
  DevArray<double> l1_scratch_, l1_scratch2_, l1_scratch3_, l1_scratch4_;
  // permutation tests
  static constexpr long long kPermScratchBytes = 1LL << 30;
  int dcc_perm_jt_ = 16, dcc_perm_rb_ = 8;
  bool dcc_shape_forced_ = false;
  hipEvent_t tev_[3] = {nullptr, nullptr, nullptr};
  DevArray<double> comp_rows_, perm_vals_, perm_obs_, perm_state_;
  DevArray<unsigned short> perm_scratch_;
  DevArray<kpc::PermUnit> d_perm_units_;
  // R's global stream (KPC_PERM_R_STREAM): the stream, the tables drawn for
  // the current wave and each pair's table offset
  bool rmode_ = false;
  kpc::RStream rs_;
  std::vector<unsigned short> rtab_host_;
  std::vector<long long> pair_rtab_;
  std::vector<int> pool_;
  // process-wide result / target-state caches (use_cache)
  bool use_results_ = false, use_targets_ = false;
  kpc::Digest dataset_;
  uint64_t opt_hash_ = 0;
  long long result_hits_ = 0, target_hits_ = 0;
  DevArray<double> d_fixed_, l1_fixed_;
  double stage_[5] = {0, 0, 0, 0, 0};
  long long launches_[5] = {0, 0, 0, 0, 0};
  double stage_host_ = 0.0;
  bool use_graph_ = std::getenv("KPC_GRAPH") != nullptr && std::atoi(std::getenv("KPC_GRAPH")) != 0;
  double graph_seconds_ = 0.0;
  long long graph_launches_ = 0;
  bool pending_smooth_checks_ = false;
};

uint64_t mask_of(const int* S, int len) {
  uint64_t m = 0;
  for (int i = 0; i < len; ++i) m |= 1ULL << S[i];
  return m;
}

bool next_comb(int* idx, int k, int len) {
  int i = k - 1;
  while (i >= 0 && idx[i] == len - k + i) --i;
  if (i < 0) return false;
  ++idx[i];
  for (int j = i + 1; j < k; ++j) idx[j] = idx[j - 1] + 1;
  return true;
}

struct Orientation {
  int x, y;
  int first = 0, count = 0;
  int nbrs = 0;
};

struct EdgeRun {
  int a, b;
  Orientation o[2];
  int side = 0, cursor = 0;
  int state = 0;  // 0 active, 1 deleted, 2 kept
};

int fill_error(const std::exception& e) {
  g_last_error = e.what();
  if (auto* k = dynamic_cast<const KpcError*>(&e)) return k->code;
  return -100;
}

kpc_result* new_result(int p) {
  kpc_result* r = (kpc_result*)std::calloc(1, sizeof(kpc_result));
  r->p = p;
  r->adjacency = (unsigned char*)std::malloc((size_t)p * p);
  r->pmax = (double*)std::malloc(sizeof(double) * (size_t)p * p);
  r->sepset_len = (int*)std::malloc(sizeof(int) * (size_t)p * p);
  r->sepset = (int*)std::calloc((size_t)p * p * 64, sizeof(int));
  for (int i = 0; i < p * p; ++i) {
    r->adjacency[i] = 1;
    r->pmax[i] = -INFINITY;
    r->sepset_len[i] = -1;
  }
  for (int i = 0; i < p; ++i) {
    r->adjacency[i * p + i] = 0;
    r->pmax[i * p + i] = 1.0;
  }
  return r;
}

void push_trace(kpc_result* r, int& cap, const Task& t, int level, int plan_index) {
  if (r->n_tasks == cap) {
    cap = cap ? cap * 2 : 4096;
    r->task_level = (int*)std::realloc(r->task_level, sizeof(int) * cap);
    r->task_x = (int*)std::realloc(r->task_x, sizeof(int) * cap);
    r->task_y = (int*)std::realloc(r->task_y, sizeof(int) * cap);
    r->task_slen = (int*)std::realloc(r->task_slen, sizeof(int) * cap);
    r->task_S = (int*)std::realloc(r->task_S, sizeof(int) * (size_t)cap * 64);
    r->task_p = (double*)std::realloc(r->task_p, sizeof(double) * cap);
    r->task_deleted = (unsigned char*)std::realloc(r->task_deleted, cap);
    r->task_margin = (double*)std::realloc(r->task_margin, sizeof(double) * cap);
    r->task_stat = (double*)std::realloc(r->task_stat, sizeof(double) * cap);
    r->task_mean = (double*)std::realloc(r->task_mean, sizeof(double) * cap);
    r->task_var = (double*)std::realloc(r->task_var, sizeof(double) * cap);
    r->task_index = (int*)std::realloc(r->task_index, sizeof(int) * cap);
    r->task_perm_seed =
        (unsigned long long*)std::realloc(r->task_perm_seed, sizeof(unsigned long long) * cap);
  }
  const int i = r->n_tasks++;
  r->task_level[i] = level;
  r->task_x[i] = t.x;
  r->task_y[i] = t.y;
  r->task_slen[i] = t.slen;
  for (int j = 0; j < 64; ++j) r->task_S[(size_t)i * 64 + j] = j < t.slen ? t.S[j] : 0;
  r->task_p[i] = t.p;
  r->task_deleted[i] = t.deleted ? 1 : 0;
  r->task_margin[i] = t.margin;
  r->task_stat[i] = t.stat;
  r->task_mean[i] = t.mean;
  r->task_var[i] = t.var;
  r->task_index[i] = plan_index + 1;
}

void run_skeleton(Engine& eng, const kpc_options& opt, kpc_result* r) {
  const int p = eng.p();
  const int mmax = opt.max_conditioning_size < 0 ? p - 2 : std::min(opt.max_conditioning_size, p - 2);
  std::vector<unsigned char> snap((size_t)p * p);
  int cap = 0;
  bool done = false;
  int ord = 0;
  const bool plain = opt.mode == KPC_PLAIN;
  long long plain_tests = 0;
  (void)plain_tests;
  // level 1's smooths, designs and fits, built on lane 1 while level 0 runs
  if (!plain && mmax >= 1) eng.prefetch_ahead();
  while (!done && ord <= mmax) {
    bool any = false;
    for (int i = 0; i < p * p; ++i) any |= r->adjacency[i] != 0;
    if (!any) break;
    done = true;
    std::memcpy(snap.data(), r->adjacency, (size_t)p * p);
    // the level's time includes building its plan, as the reference's
    // level_started precedes make_layer_plan
    const double level_t0 = now_s();
    // ---- level plan
    std::vector<Task> tasks;
    std::vector<EdgeRun> edges;
    int nb[64], idx[64];
    {
      // the plan's size first (sum over orientations of C(neighbours, ord)),
      // so the task vector is allocated once
      double count = 0.0;
      for (int a = 0; a < p; ++a) {
        int deg = 0;
        for (int i = 0; i < p; ++i) deg += i != a && snap[(size_t)i * p + a];
        if (deg - 1 < ord) continue;
        double c = 1.0;
        for (int i = 0; i < ord; ++i) c = c * (double)(deg - 1 - i) / (double)(i + 1);
        count += deg * c;
      }
      if (count < 1e7) tasks.reserve((size_t)count);
    }
    for (int a = 0; a < p - 1; ++a)
      for (int b = a + 1; b < p; ++b) {
        if (!snap[(size_t)a * p + b]) continue;
        EdgeRun er;
        er.a = a;
        er.b = b;
        for (int side = 0; side < 2; ++side) {
          const int x = side ? b : a, y = side ? a : b;
          Orientation& o = er.o[side];
          o.x = x;
          o.y = y;
          o.first = (int)tasks.size();
          int len = 0;
          for (int i = 0; i < p; ++i)
            if (i != y && snap[(size_t)i * p + x]) nb[len++] = i;
          o.nbrs = len;
          if (len >= ord) {
            for (int i = 0; i < ord; ++i) idx[i] = i;
            do {
              Task t;
              t.x = x;
              t.y = y;
              t.slen = ord;
              for (int i = 0; i < ord; ++i) t.S[i] = nb[idx[i]];
              t.mask = mask_of(t.S, ord);
              tasks.push_back(t);
            } while (ord > 0 && next_comb(idx, ord, len));
          }
          o.count = (int)tasks.size() - o.first;
        }
        edges.push_back(er);
      }
    double stage0[5];
    for (int st = 0; st < 5; ++st) stage0[st] = eng.stage(st);
    // Levels whose whole plan needs at most this many residual fits are
    // evaluated speculatively in one wave (fewer round trips beat the wasted
    // fits); larger levels run frontier waves with doubling chunks.
    size_t level_unique_resid = 0;
    size_t speculate_limit = 16384;
    if (const char* env = std::getenv("KPC_SPECULATE_LIMIT")) speculate_limit = std::atol(env);
    // the next level is built ahead when its superset needs at most this many
    // fits (KPC_AHEAD_LIMIT, 0 turns it off)
    size_t ahead_limit = 512;
    if (const char* env = std::getenv("KPC_AHEAD_LIMIT")) ahead_limit = std::atol(env);
    const Engine::Stats st0 = eng.stats_;
    eng.begin_level(ord);
    {
      FlatMap<uint64_t, int, U64Hash> ud;
      FlatMap<std::pair<int, uint64_t>, int, KeyHash> ur;
      ud.reserve(tasks.size());
      ur.reserve(2 * tasks.size());
      for (const Task& t : tasks) {
        ud.emplace(t.mask, 0);
        ur.emplace(std::make_pair(t.x, t.mask), 0);
        ur.emplace(std::make_pair(t.y, t.mask), 0);
      }
      eng.reserve_level(ud.size(), ur.size());
      level_unique_resid = ur.size();
      if (!plain) eng.adopt_ahead(ud.size(), ur.size());
    }
    if (ord == 1 && !plain) eng.adopt_level1();
    // the next level's designs and fits, built ahead from this level's graph
    // behind this level's own GCV work (a level that claimed its fits from
    // such a chain is too short to hide another one: if the skeleton ended
    // there, its end would wait for the whole chain)
    if (!plain && ord + 1 <= mmax && !eng.claims_ahead())
      eng.plan_ahead(snap.data(), ord + 1, ahead_limit);
    if (eng.r_stream()) {
      // R-stream permutations: the tests run in pcalg's execution order
      // (orientations (x, y) by x, then y; each orientation's conditioning
      // sets in lexicographic order; fastkpc/src/full_cuda_ci_one_call.cpp:
      // 5301-5330 with pcalg_task_less :326-336) and each draws its
      // replicates from the stream as it runs.  Every fit and component of the
      // level is computed first in one wave; then a wave takes the next
      // `chunk` tests of one orientation, drawn as if all of them ran, and the
      // stream is rewound to the end of the last test consumed (a deletion
      // ends the orientation and drops the tests drawn after it).
      std::vector<Task*> all;
      all.reserve(tasks.size());
      for (Task& t : tasks) all.push_back(&t);
      eng.evaluate(all, false);
      ++r->level_rounds[ord];
      std::vector<std::pair<int, int>> order;  // (edge, side)
      for (size_t ei = 0; ei < edges.size(); ++ei)
        for (int side = 0; side < 2; ++side) order.emplace_back((int)ei, side);
      std::sort(order.begin(), order.end(),
                [&](const std::pair<int, int>& u, const std::pair<int, int>& v) {
                  const Orientation& a = edges[u.first].o[u.second];
                  const Orientation& b = edges[v.first].o[v.second];
                  return a.x != b.x ? a.x < b.x : a.y < b.y;
                });
      std::vector<Task*> batch;
      std::vector<kpc::RStream::State> snaps;
      for (const auto& os : order) {
        EdgeRun& e = edges[os.first];
        if (e.state == 1) continue;  // deleted by its other orientation
        const Orientation& o = e.o[os.second];
        if (o.count == 0) continue;  // fewer neighbours than the level
        if (o.nbrs > ord) done = false;
        int c = 0, chunk = 4;
        bool deleted = false;
        while (c < o.count && !deleted) {
          const int take = std::min(chunk, o.count - c);
          batch.clear();
          snaps.clear();
          eng.clear_tables();
          for (int j = 0; j < take; ++j) {
            Task& t = tasks[(size_t)o.first + c + j];
            snaps.push_back(eng.rstate());
            eng.draw_table(t);
            batch.push_back(&t);
          }
          snaps.push_back(eng.rstate());
          eng.evaluate(batch);
          ++r->level_rounds[ord];
          int used = take;
          for (int j = 0; j < take; ++j)
            if (batch[j]->p >= opt.alpha) {
              used = j + 1;
              deleted = true;
              break;
            }
          eng.set_rstate(snaps[used]);
          for (int j = 0; j < used; ++j) {
            Task& t = *batch[j];
            t.consumed = true;
            ++r->n_edgetests[ord];
            double& pm = r->pmax[(size_t)t.x * p + t.y];
            if (pm < t.p) pm = t.p;
          }
          for (int j = used; j < take; ++j) batch[j]->computed = false;  // drawn, not run
          if (deleted) {
            const Task& t = *batch[used - 1];
            batch[used - 1]->deleted = true;
            e.state = 1;
            e.side = os.second;
            e.cursor = c + used - 1;
            r->adjacency[(size_t)e.a * p + e.b] = 0;  // applied after the level via snap
            r->sepset_len[(size_t)e.a * p + e.b] = r->sepset_len[(size_t)e.b * p + e.a] = ord;
            for (int i = 0; i < ord; ++i) {
              r->sepset[((size_t)e.a * p + e.b) * 64 + i] = t.S[i];
              r->sepset[((size_t)e.b * p + e.a) * 64 + i] = t.S[i];
            }
          }
          c += used;
          chunk = std::min(chunk * 2, 1 << 12);
        }
      }
    } else {
      // ---- frontier waves
      for (EdgeRun& e : edges) {
        e.side = 0;
        e.cursor = 0;
        if (e.o[0].nbrs > ord) done = false;
        if (e.o[0].count == 0) {
          e.side = 1;
          if (e.o[1].nbrs > ord) done = false;
          if (e.o[1].count == 0) e.state = 2;
        }
      }
      // frontier chunk: tasks taken per undecided orientation in the first
      // wave, doubled every wave (KPC_WAVE_CHUNK overrides; <= 0 evaluates the
      // whole level speculatively in one wave)
      int chunk = level_unique_resid <= speculate_limit ? 0 : 4;
      if (const char* env = std::getenv("KPC_WAVE_CHUNK")) chunk = std::atoi(env);
      if (chunk <= 0) chunk = 1 << 30;
      const int batched_chunk = chunk;
      bool plain_now = plain;
      if (plain_now) chunk = 1;
      bool speculate_all = chunk >= (1 << 30);
      long long level_plain = 0;
      bool active = true;
      std::vector<Task*> batch;
      while (active) {
        batch.clear();
        for (EdgeRun& e : edges) {
          if (e.state != 0) continue;
          for (int side = e.side; side < 2; ++side) {
            const Orientation& o = e.o[side];
            int taken = 0;
            for (int c = side == e.side ? e.cursor : 0; c < o.count && taken < chunk; ++c) {
              Task& t = tasks[(size_t)o.first + c];
              if (!t.computed) {
                batch.push_back(&t);
                ++taken;
              }
            }
            if (!speculate_all) break;
          }
          if (plain_now && !batch.empty()) break;
        }
        ++r->level_rounds[ord];
        if (plain_now) {
          // one launch sequence per test: fresh caches for every test
          eng.set_share_univariate(false);
          eng.begin_level(ord);
          const double tp0 = now_s();
          eng.evaluate(batch);
          r->plain_seconds[ord] += now_s() - tp0;
          r->plain_tests[ord] += (long long)batch.size();
          ++level_plain;
          plain_tests += (long long)batch.size();
          if (opt.plain_test_limit > 0 && level_plain >= opt.plain_test_limit) {
            plain_now = false;  // rest of the level runs batched
            chunk = batched_chunk;
            speculate_all = chunk >= (1 << 30);
            eng.set_share_univariate(true);
            eng.begin_level(ord);
          }
        } else {
          eng.evaluate(batch);
        }
        // replay in canonical order
        active = false;
        for (EdgeRun& e : edges) {
          while (e.state == 0) {
            const Orientation& o = e.o[e.side];
            if (e.cursor < o.count) {
              Task& t = tasks[(size_t)o.first + e.cursor];
              if (!t.computed) break;
              t.consumed = true;
              ++r->n_edgetests[ord];
              double& pm = r->pmax[(size_t)t.x * p + t.y];
              if (pm < t.p) pm = t.p;
              if (t.p >= opt.alpha) {
                t.deleted = true;
                e.state = 1;
                r->adjacency[(size_t)e.a * p + e.b] = 0;  // applied after the level via snap
                r->sepset_len[(size_t)e.a * p + e.b] = r->sepset_len[(size_t)e.b * p + e.a] = ord;
                for (int i = 0; i < ord; ++i) {
                  r->sepset[((size_t)e.a * p + e.b) * 64 + i] = t.S[i];
                  r->sepset[((size_t)e.b * p + e.a) * 64 + i] = t.S[i];
                }
                break;
              }
              ++e.cursor;
            } else if (e.side == 0) {
              e.side = 1;
              e.cursor = 0;
              if (e.o[1].nbrs > ord) done = false;
              if (e.o[1].count == 0) e.state = 2;
            } else {
              e.state = 2;
            }
          }
          if (e.state == 0) active = true;
        }
        if (!plain) chunk = std::min(chunk * 2, 1 << 20);
      }
    }
    for (int a = 0; a < p; ++a)
      for (int b = 0; b < p; ++b)
        if (!r->adjacency[(size_t)a * p + b] || !r->adjacency[(size_t)b * p + a]) {
          r->adjacency[(size_t)a * p + b] = 0;
          r->adjacency[(size_t)b * p + a] = 0;
        }
    // tasks left on an edge after the test that deleted it (both
    // orientations' tasks form the edge's run in the plan)
    for (const EdgeRun& e : edges) {
      if (e.state != 1) continue;
      const int total = e.o[0].count + e.o[1].count;
      const int pos = (e.side ? e.o[0].count : 0) + e.cursor;  // the deleting task
      r->level_tasks_ignored[ord] += total - (pos + 1);
    }
    const bool perm = opt.ci_method == KPC_DCC_PERM || opt.ci_method == KPC_HSIC_PERM;
    const uint64_t user_seed = opt.has_seed ? (uint64_t)opt.seed : 0ULL;
    for (size_t ti = 0; ti < tasks.size(); ++ti) {
      if (!tasks[ti].consumed) continue;
      push_trace(r, cap, tasks[ti], ord, (int)ti);
      const Task& t = tasks[ti];
      r->task_perm_seed[r->n_tasks - 1] =
          perm ? task_seed(user_seed, opt.ci_method, t.x, t.y, t.S, t.slen) : 0ULL;
    }
    r->level_tests[ord] = r->n_edgetests[ord];
    r->level_tasks_planned[ord] = (long long)tasks.size();
    r->level_physical[ord] = eng.stats_.physical_tests - st0.physical_tests;
    r->level_fits[ord] = eng.stats_.fits - st0.fits;
    r->level_components[ord] = eng.stats_.components - st0.components;
    r->level_evaluations[ord] = eng.stats_.evals - st0.evals;
    r->level_lanczos_steps[ord] = eng.stats_.lanczos_steps - st0.lanczos_steps;
    r->level_seconds[ord] = now_s() - level_t0;
    for (int st = 0; st < 5; ++st) r->level_stage_seconds[ord][st] = eng.stage(st) - stage0[st];
    ++ord;
  }
  r->levels = ord;
  {
    std::string keys;
    keys.reserve((size_t)r->n_tasks * 8 + 1);
    char num[12];
    for (int i = 0; i < r->n_tasks; ++i) {
      if (i) keys.push_back('\n');
      for (int j = 0; j < r->task_slen[i]; ++j) {
        if (j) keys.push_back('|');
        unsigned v = (unsigned)(r->task_S[(size_t)i * 64 + j] + 1);
        int len = 0;
        do {
          num[len++] = (char)('0' + v % 10);
          v /= 10;
        } while (v);
        while (len) keys.push_back(num[--len]);
      }
    }
    r->task_keys = (char*)std::malloc(keys.size() + 1);
    std::memcpy(r->task_keys, keys.c_str(), keys.size() + 1);
  }
  for (int i = 0; i < p - 1; ++i)
    for (int j = i + 1; j < p; ++j) {
      const double m = std::max(r->pmax[(size_t)i * p + j], r->pmax[(size_t)j * p + i]);
      r->pmax[(size_t)i * p + j] = r->pmax[(size_t)j * p + i] = m;
    }
}

}  // namespace

extern "C" {

void kpc_default_options(kpc_options* o) {
  std::memset(o, 0, sizeof(*o));
  o->ci_method = KPC_DCC_GAMMA;
  o->alpha = 0.1;
  o->max_conditioning_size = -1;
  o->index = 1.0;
  o->num_col = 35;
  o->hsic_sig = 1.0;
  o->replicates = 100;
  o->include_observed = 1;
  o->mode = KPC_BATCHED;
}

int kpc_skeleton_run(const double* data, int n, int p, const kpc_options* opt, kpc_result** out) {
  try {
    if (!data || !opt || !out) throw KpcError(-1, "null argument");
    if (!(opt->alpha > 0.0 && opt->alpha < 1.0)) throw KpcError(-1, "alpha must be in (0, 1)");
    const double t0 = now_s();
    Engine eng(data, n, p, *opt);
    kpc_result* r = new_result(p);
    try {
      run_skeleton(eng, *opt, r);
    } catch (...) {
      kpc_result_free(r);
      throw;
    }
    r->wall_seconds = now_s() - t0;
    for (int s = 0; s < 5; ++s) {
      r->stage_seconds[s] = eng.stage(s);
      r->stage_launches[s] = eng.launches(s);
    }
    r->stage_seconds[5] = eng.stage_host();
    r->stage_seconds[7] = eng.graph_seconds();
    r->stage_launches[7] = eng.graph_launches();
    if (eng.r_stream()) eng.save_rstate(r->r_state_out);
    r->result_cache_hits = eng.result_hits();
    r->target_cache_hits = eng.target_hits();
    r->physical_tests = eng.stats_.physical_tests;
    r->residual_fits = eng.stats_.fits;
    r->components = eng.stats_.components;
    r->setups = eng.stats_.setups;
    r->magic_evaluations = eng.stats_.evals;
    *out = r;
    return 0;
  } catch (const std::exception& e) {
    return fill_error(e);
  }
}

void kpc_result_free(kpc_result* r) {
  if (!r) return;
  std::free(r->adjacency);
  std::free(r->pmax);
  std::free(r->sepset_len);
  std::free(r->sepset);
  std::free(r->task_level);
  std::free(r->task_x);
  std::free(r->task_y);
  std::free(r->task_slen);
  std::free(r->task_S);
  std::free(r->task_p);
  std::free(r->task_deleted);
  std::free(r->task_margin);
  std::free(r->task_keys);
  std::free(r->task_stat);
  std::free(r->task_mean);
  std::free(r->task_var);
  std::free(r->task_index);
  std::free(r->task_perm_seed);
  std::free(r);
}

int kpc_ci_batch(const double* data, int n, int p, const kpc_options* opt, int count, const int* x,
                 const int* y, int s_len, const int* S, double* out) {
  return kpc_ci_batch_rs(data, n, p, opt, count, x, y, s_len, S, out, nullptr);
}

int kpc_ci_batch_rs(const double* data, int n, int p, const kpc_options* opt, int count,
                    const int* x, const int* y, int s_len, const int* S, double* out,
                    int* r_state_out) {
  try {
    if (!data || !opt || !out || count < 0 || s_len < 0 || s_len > 62)
      throw KpcError(-1, "invalid arguments");
    Engine eng(data, n, p, *opt);
    eng.begin_level(s_len);
    std::vector<Task> tasks((size_t)count);
    for (int i = 0; i < count; ++i) {
      Task& t = tasks[i];
      t.x = x[i];
      t.y = y[i];
      t.slen = s_len;
      if (t.x < 0 || t.x >= p || t.y < 0 || t.y >= p || t.x == t.y)
        throw KpcError(-1, "x/y out of range");
      for (int j = 0; j < s_len; ++j) {
        t.S[j] = S[(size_t)i * s_len + j];
        if (t.S[j] < 0 || t.S[j] >= p || t.S[j] == t.x || t.S[j] == t.y)
          throw KpcError(-1, "S out of range or overlaps x/y");
        if (j > 0 && t.S[j] <= t.S[j - 1]) throw KpcError(-1, "S must be strictly increasing");
      }
      t.mask = mask_of(t.S, s_len);
    }
    eng.reserve_level((size_t)count, (size_t)count * 2);
    std::vector<Task*> ptrs;
    for (Task& t : tasks) {
      if (eng.r_stream()) eng.draw_table(t);  // batch order = draw order
      ptrs.push_back(&t);
    }
    eng.evaluate(ptrs);
    if (eng.r_stream() && r_state_out) eng.save_rstate(r_state_out);
    for (int i = 0; i < count; ++i) {
      out[4 * i + 0] = tasks[i].p;
      out[4 * i + 1] = tasks[i].stat;
      out[4 * i + 2] = tasks[i].mean;
      out[4 * i + 3] = tasks[i].var;
    }
    return 0;
  } catch (const std::exception& e) {
    return fill_error(e);
  }
}

namespace {

// Conditioning sets of one |S| group, keyed by id (distinct sorted S lists
// share a design), validated against p.
std::vector<Task> fit_sets(int p, int count, const int* targets, int s_len,
                           const int* const* S_rows) {
  std::vector<Task> sets((size_t)count);
  std::map<std::vector<int>, uint64_t> ids;
  for (int i = 0; i < count; ++i) {
    Task& t = sets[i];
    t.slen = s_len;
    t.x = targets[i];
    if (targets[i] < 0 || targets[i] >= p) throw KpcError(-1, "target out of range");
    std::vector<int> key((size_t)s_len);
    for (int j = 0; j < s_len; ++j) {
      t.S[j] = S_rows[i][j];
      if (t.S[j] < 0 || t.S[j] >= p) throw KpcError(-1, "S out of range");
      // any order of distinct columns: the order of the smooth terms in the
      // GAM formula (kpcalg's regrVonPS regresses on c(S, parents))
      for (int i = 0; i < j; ++i)
        if (t.S[i] == t.S[j]) throw KpcError(-1, "S must hold distinct columns");
      if (t.S[j] == t.x) throw KpcError(-1, "target must not be in S");
      key[j] = t.S[j];
    }
    auto it = ids.emplace(key, (uint64_t)ids.size()).first;
    t.mask = it->second;
  }
  return sets;
}

// HIP-event seconds of the last regrXonS call on this thread:
// {design setup (smooth + geometry), GCV fits, launches of the GCV kernel}.
thread_local double g_fit_timing[3] = {0.0, 0.0, 0.0};

// GCV fits of `count` targets sharing |S| == s_len on an engine whose data
// is resident: designs (smooth setups + geometry), then one magic launch.
// resid / info rows go to out_index[i] (count x n, count x 32).
void run_fit_group(Engine& eng, int count, const int* targets, int s_len,
                   const int* const* S_rows, const int* out_index, double* resid, double* info) {
  const int n = eng.n(), p = eng.p();
  eng.begin_level(s_len);
  std::vector<Task> sets = fit_sets(p, count, targets, s_len, S_rows);
  eng.reserve_level((size_t)count, (size_t)count);
  hipEvent_t ev[3];
  for (auto& e : ev) hip_check(hipEventCreate(&e), "hipEventCreate");
  hip_check(hipEventRecord(ev[0], eng.stream()), "event");
  std::vector<int> dslots = eng.build_designs(sets);
  std::vector<kpc::MagicItem> items((size_t)count);
  for (int i = 0; i < count; ++i) {
    items[i].design_slot = dslots[i];
    items[i].target = targets[i];
    items[i].resid_slot = i;
    items[i].pad = 0;
  }
  DevArray<kpc::MagicItem> d_items;
  d_items.upload(items, eng.stream());
  const long long stride = kpc::magic_scratch_doubles(eng.q(), eng.m(), eng.rk(), n);
  const int grid = std::min(count, eng.slab_grid(stride, 8192));
  eng.scratch3().reserve((size_t)grid * stride);
  hip_check(hipEventRecord(ev[1], eng.stream()), "event");
  hip_check(kpc::launch_magic(d_items.ptr, count, n, eng.q(), eng.m(), eng.rk(), eng.data_ptr(),
                              eng.design_ptr(), eng.resid_arena().ptr, eng.rinfo_arena().ptr,
                              eng.scratch3().ptr, grid, eng.ctl(), eng.stream()),
            "magic launch");
  hip_check(hipEventRecord(ev[2], eng.stream()), "event");
  hip_check(hipStreamSynchronize(eng.stream()), "magic synchronize");
  float ms0 = 0.f, ms1 = 0.f;
  hip_check(hipEventElapsedTime(&ms0, ev[0], ev[1]), "event elapsed");
  hip_check(hipEventElapsedTime(&ms1, ev[1], ev[2]), "event elapsed");
  g_fit_timing[0] += ms0 * 1e-3;
  g_fit_timing[1] += ms1 * 1e-3;
  g_fit_timing[2] += 1.0;
  for (auto& e : ev) (void)hipEventDestroy(e);
  std::vector<double> r((size_t)count * n), f((size_t)count * kpc::kMagicInfo);
  hip_check(hipMemcpy(r.data(), eng.resid_arena().ptr, sizeof(double) * r.size(),
                      hipMemcpyDeviceToHost),
            "D2H residuals");
  hip_check(hipMemcpy(f.data(), eng.rinfo_arena().ptr, sizeof(double) * f.size(),
                      hipMemcpyDeviceToHost),
            "D2H fit info");
  for (int i = 0; i < count; ++i) {
    const size_t o = (size_t)out_index[i];
    std::memcpy(resid + o * n, r.data() + (size_t)i * n, sizeof(double) * n);
    if (info)
      std::memcpy(info + o * kpc::kMagicInfo, f.data() + (size_t)i * kpc::kMagicInfo,
                  sizeof(double) * kpc::kMagicInfo);
  }
}

Engine make_fit_engine(const double* data, int n, int p) {
  kpc_options o;
  kpc_default_options(&o);
  o.num_col = std::min(35, n - 1);
  return Engine(data, n, p, o, false);
}

}  // namespace

int kpc_regrxons_batch(const double* data, int n, int p, int count, const int* targets, int s_len,
                       const int* S, double* resid, double* info) {
  try {
    if (!data || !targets || !S || !resid || count <= 0 || s_len < 1 || s_len > 62)
      throw KpcError(-1, "invalid arguments");
    Engine eng = make_fit_engine(data, n, p);
    g_fit_timing[0] = g_fit_timing[1] = g_fit_timing[2] = 0.0;
    std::vector<const int*> rows((size_t)count);
    std::vector<int> order((size_t)count);
    for (int i = 0; i < count; ++i) {
      rows[i] = S + (size_t)i * s_len;
      order[i] = i;
    }
    run_fit_group(eng, count, targets, s_len, rows.data(), order.data(), resid, info);
    return 0;
  } catch (const std::exception& e) {
    return fill_error(e);
  }
}

int kpc_regrxons_mixed(const double* data, int n, int p, int count, const int* targets,
                       const int* s_len, const int* s_off, const int* S, double* resid,
                       double* info) {
  try {
    if (!data || !targets || !s_len || !s_off || !S || !resid || count <= 0)
      throw KpcError(-1, "invalid arguments");
    for (int i = 0; i < count; ++i)
      if (s_len[i] < 1 || s_len[i] > 62 || s_off[i] < 0)
        throw KpcError(-1, "each |S| must be in [1, 62] with a non-negative offset");
    Engine eng = make_fit_engine(data, n, p);
    g_fit_timing[0] = g_fit_timing[1] = g_fit_timing[2] = 0.0;
    for (int L = 1; L <= 62; ++L) {
      std::vector<int> idx, tg;
      std::vector<const int*> rows;
      for (int i = 0; i < count; ++i)
        if (s_len[i] == L) {
          idx.push_back(i);
          tg.push_back(targets[i]);
          rows.push_back(S + s_off[i]);
        }
      if (idx.empty()) continue;
      run_fit_group(eng, (int)idx.size(), tg.data(), L, rows.data(), idx.data(), resid, info);
    }
    return 0;
  } catch (const std::exception& e) {
    return fill_error(e);
  }
}

int kpc_tuning_report(char* buf, int cap) {
  if (!buf || cap <= 0) return -1;
  const std::string r = kpc::Tuner::instance().report();
  std::snprintf(buf, (size_t)cap, "%s", r.c_str());
  return (int)r.size();
}

int kpc_last_fit_timing(double* out) {
  if (!out) return -1;
  for (int i = 0; i < 3; ++i) out[i] = g_fit_timing[i];
  return 0;
}

int kpc_gcv_grad_hess(const double* data, int n, int p, int count, const int* targets, int s_len,
                      const int* S, const double* log_sp, double* score, double* grad,
                      double* hess) {
  try {
    if (!data || !targets || !S || !log_sp || count <= 0 || s_len < 1 || s_len > 62)
      throw KpcError(-1, "invalid arguments");
    Engine eng = make_fit_engine(data, n, p);
    eng.begin_level(s_len);
    std::vector<const int*> rows((size_t)count);
    for (int i = 0; i < count; ++i) rows[i] = S + (size_t)i * s_len;
    std::vector<Task> sets = fit_sets(p, count, targets, s_len, rows.data());
    eng.reserve_level((size_t)count, (size_t)count);
    std::vector<int> dslots = eng.build_designs(sets);
    const int m = eng.m();
    std::vector<kpc::MagicItem> items((size_t)count);
    for (int i = 0; i < count; ++i) {
      items[i].design_slot = dslots[i];
      items[i].target = targets[i];
      items[i].resid_slot = i;
      items[i].pad = 0;
    }
    DevArray<kpc::MagicItem> d_items;
    d_items.upload(items, eng.stream());
    std::vector<double> hl(log_sp, log_sp + (size_t)count * m);
    DevArray<double> d_lsp, d_out;
    d_lsp.upload(hl, eng.stream());
    const size_t per = (size_t)(1 + m + m * m);
    d_out.reserve(per * count);
    const long long stride = kpc::magic_scratch_doubles(eng.q(), m, eng.rk(), n);
    const int grid = std::min(count, eng.slab_grid(stride, 8192));
    eng.scratch3().reserve((size_t)grid * stride);
    hip_check(kpc::launch_gcv_eval(d_items.ptr, count, n, eng.q(), m, eng.rk(), eng.data_ptr(),
                                   eng.design_ptr(), d_lsp.ptr, d_out.ptr, eng.scratch3().ptr, grid,
                                   eng.ctl().rank_tol, eng.stream()),
              "gcv eval launch");
    hip_check(hipStreamSynchronize(eng.stream()), "gcv synchronize");
    std::vector<double> h(per * count);
    hip_check(hipMemcpy(h.data(), d_out.ptr, sizeof(double) * h.size(), hipMemcpyDeviceToHost),
              "D2H gcv");
    for (int i = 0; i < count; ++i) {
      const double* o2 = h.data() + per * i;
      if (score) score[i] = o2[0];
      if (grad) for (int j = 0; j < m; ++j) grad[(size_t)i * m + j] = o2[1 + j];
      if (hess) for (int j = 0; j < m * m; ++j) hess[(size_t)i * m * m + j] = o2[1 + m + j];
    }
    return 0;
  } catch (const std::exception& e) {
    return fill_error(e);
  }
}

int kpc_dcov_components(const double* x, int n, int count, int k, double* values, double* vectors,
                        double* dsum, double* moment) {
  try {
    if (!x || count <= 0 || n < 6 || n > kpc::kMaxN || k <= 0 || k >= n || k > kpc::kMaxNumCol)
      throw KpcError(-1, "invalid arguments");
    for (long long i = 0; i < (long long)n * count; ++i)
      if (!std::isfinite(x[i])) throw KpcError(-1, "data must be finite");
    DevArray<double> d_x, d_vals, d_vecs, d_dsum, d_mom, d_scr;
    DevArray<long long> d_src;
    DevArray<int> d_dst, d_st;
    StreamGuard guard;
    hipStream_t s = guard.s;
    std::vector<double> hx(x, x + (size_t)n * count);
    d_x.upload(hx, s);
    std::vector<long long> src(count);
    std::vector<int> dst(count);
    for (int i = 0; i < count; ++i) { src[i] = (long long)i * n; dst[i] = i; }
    d_src.upload(src, s);
    d_dst.upload(dst, s);
    d_vals.reserve((size_t)count * k);
    d_vecs.reserve((size_t)count * n * k);
    d_dsum.reserve(count);
    d_mom.reserve(count);
    d_st.reserve(count);
    if (std::getenv("KPC_DCOV_DENSE") != nullptr) {
      const long long stride = kpc::dcov_component_scratch_doubles(n, k);
      const int grid = std::min(count, 1024);
      d_scr.reserve((size_t)grid * stride);
      hip_check(kpc::launch_dcov_components(d_x.ptr, d_src.ptr, d_dst.ptr, count, n, k, d_vals.ptr,
                                            d_vecs.ptr, d_dsum.ptr, d_mom.ptr, d_scr.ptr, grid, s),
                "component launch");
    } else {
      const int mmax = kpc::lanczos_mmax(n);
      const long long stride = kpc::dcov_lanczos_scratch_doubles(n, k, mmax);
      const int grid = std::min(count, 1024);
      d_scr.reserve((size_t)grid * stride);
      hip_check(kpc::launch_dcov_lanczos(d_x.ptr, d_src.ptr, d_dst.ptr, count, n, k, d_vals.ptr,
                                         d_vecs.ptr, d_dsum.ptr, d_mom.ptr, d_st.ptr, d_scr.ptr,
                                         grid, mmax, 1e-13, s),
                "lanczos launch");
    }
    hip_check(hipStreamSynchronize(s), "component synchronize");
    if (values) hip_check(hipMemcpy(values, d_vals.ptr, sizeof(double) * count * k, hipMemcpyDeviceToHost), "D2H");
    if (vectors) hip_check(hipMemcpy(vectors, d_vecs.ptr, sizeof(double) * (size_t)count * n * k, hipMemcpyDeviceToHost), "D2H");
    if (dsum) hip_check(hipMemcpy(dsum, d_dsum.ptr, sizeof(double) * count, hipMemcpyDeviceToHost), "D2H");
    if (moment) hip_check(hipMemcpy(moment, d_mom.ptr, sizeof(double) * count, hipMemcpyDeviceToHost), "D2H");
    return 0;
  } catch (const std::exception& e) {
    return fill_error(e);
  }
}

namespace {
void fill_cache_info(kpc_cache_info* out, const kpc::Digest* dataset) {
  const kpc::CacheStats a = kpc::result_cache().stats(dataset);
  const kpc::CacheStats b = kpc::target_cache().stats(dataset);
  std::memset(out, 0, sizeof(*out));
  out->capacity = a.capacity;
  out->entries = a.entries;
  out->dataset_entries = a.dataset_entries;
  out->requests = a.requests;
  out->hits = a.hits;
  out->misses = a.misses;
  out->inserts = a.inserts;
  out->evictions = a.evictions;
  out->generation = a.generation;
  out->target_capacity = b.capacity;
  out->target_entries = b.entries;
  out->target_dataset_entries = b.dataset_entries;
  out->target_requests = b.requests;
  out->target_hits = b.hits;
  out->target_misses = b.misses;
  out->target_inserts = b.inserts;
  out->target_evictions = b.evictions;
  out->target_generation = b.generation;
  out->epoch = kpc::cache_epoch().load();
  if (dataset)
    std::snprintf(out->dataset_key, sizeof(out->dataset_key), "%016llx%016llx",
                  (unsigned long long)dataset->hi, (unsigned long long)dataset->lo);
}
}  // namespace

int kpc_cache_control(int action, long long capacity, kpc_cache_info* out) {
  try {
    if (!out) throw KpcError(-1, "null argument");
    switch (action) {
      case KPC_CACHE_INFO:
        if (capacity != -1) throw KpcError(-1, "cache info does not accept a capacity");
        break;
      case KPC_CACHE_RESET:
        if (capacity != -1) throw KpcError(-1, "cache reset does not accept a capacity");
        kpc::result_cache().reset();
        kpc::target_cache().reset();
        kpc::cache_epoch().fetch_add(1);
        break;
      case KPC_CACHE_CONFIGURE:
        if (!kpc::result_cache().configure(capacity))
          throw KpcError(-1, "result cache capacity is outside [1, 1048576]");
        break;
      case KPC_CACHE_CONFIGURE_TARGET:
        if (!kpc::target_cache().configure(capacity))
          throw KpcError(-1, "target-state cache capacity is outside [1, 524288]");
        break;
      default:
        throw KpcError(-1, "unknown cache control action");
    }
    fill_cache_info(out, nullptr);
    return 0;
  } catch (const std::exception& e) {
    return fill_error(e);
  }
}

int kpc_cache_state(const double* data, int n, int p, kpc_cache_info* out) {
  try {
    if (!data || !out || n <= 0 || p <= 0) throw KpcError(-1, "invalid arguments");
    for (long long i = 0; i < (long long)n * p; ++i)
      if (!std::isfinite(data[i])) throw KpcError(-1, "data must be finite");
    const kpc::Digest d = kpc::dataset_digest(data, n, p);
    fill_cache_info(out, &d);
    return 0;
  } catch (const std::exception& e) {
    return fill_error(e);
  }
}

int kpc_task_seed(unsigned long long user_seed, int ci_method, int x, int y, const int* S,
                  int s_len, unsigned long long* seed) {
  try {
    if (!seed || s_len < 0 || s_len > 64 || (s_len > 0 && !S)) throw KpcError(-1, "invalid arguments");
    *seed = task_seed(user_seed, ci_method, x, y, S, s_len);
    return 0;
  } catch (const std::exception& e) {
    return fill_error(e);
  }
}

int kpc_permutation_table(int n, int replicates, int count, const unsigned long long* seeds,
                          int* out) {
  try {
    if (!seeds || !out || n < 2 || n > kpc::kMaxN || replicates < 1 || count < 1)
      throw KpcError(-1, "invalid arguments");
    const size_t total = (size_t)count * replicates * n;
    DevArray<unsigned long long> d_seeds;
    DevArray<int> d_out;
    StreamGuard guard;
    hipStream_t s = guard.s;
    d_seeds.upload(std::vector<unsigned long long>(seeds, seeds + count), s);
    d_out.reserve(total);
    hip_check(kpc::launch_permutation_table(d_seeds.ptr, count, n, replicates, d_out.ptr, s),
              "permutation launch");
    hip_check(hipStreamSynchronize(s), "permutation synchronize");
    hip_check(hipMemcpy(out, d_out.ptr, sizeof(int) * total, hipMemcpyDeviceToHost), "D2H");
    return 0;
  } catch (const std::exception& e) {
    return fill_error(e);
  }
}

const char* kpc_last_error(void) { return g_last_error.c_str(); }

int kpc_empty_cache(void) {
  try {
    pool_empty();
    return 0;
  } catch (const std::exception& e) {
    return fill_error(e);
  }
}

int kpc_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

}  // extern "C"
