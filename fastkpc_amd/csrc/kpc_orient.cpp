// kpcalg-authority WAN-PDAG orientation on top of the library's GPU regrXonS
// and CI tests: the counterpart of fastkpc_orient_wanpdag_kpcalg_authority
// (fastkpc/R/fast_kpc.R:263-445), which runs kpcalg's udag2wanpdag
// (kpcalg/R/udag2wanpdag.R:104-436) with regrVonPS (kpcalg/R/regrvonps.R:
// 17-39) on the CPU.  The graph algorithm is host code with R's iteration
// order (which(arr.ind = TRUE) column by column, combn order, loops that do
// not stop early); each regrVonPS call is one GPU fit of V on c(S, parents)
// in that column order and one batch of unconditional CI tests of the
// residual against the raw columns of S, through the public C ABI.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "kpcmi.h"

namespace kpc {
void set_last_error(const char* msg);
}

namespace {

// lib_error: a library entry point failed and has set kpc_last_error itself
struct OrientError : std::runtime_error {
  int code;
  bool lib_error;
  OrientError(int c, const std::string& m, bool lib = false)
      : std::runtime_error(m), code(c), lib_error(lib) {}
};

uint64_t splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

class Orienter {
 public:
  Orienter(const double* data, int n, int p, const kpc_orient_options& opt, kpc_orient_result* r)
      : data_(data), n_(n), p_(p), opt_(opt), r_(r) {
    if (opt.perm_stream == KPC_PERM_R_STREAM) {
      if (opt.r_state) std::memcpy(rstate_, opt.r_state, sizeof(rstate_));
      else if (!opt.has_seed)
        throw OrientError(-1, "R-stream permutations need a seed or a stream state");
      rstate_valid_ = opt.r_state != nullptr;
    }
  }

  void run(const unsigned char* adjacency, const int* sepset_len, const int* sepset) {
    const int p = p_;
    std::vector<int> g((size_t)p * p);
    int edges = 0;
    for (int i = 0; i < p * p; ++i) edges += (g[i] = adjacency[i] ? 1 : 0);
    pdag_ = g;
    if (edges > 0) {
      if (opt_.orient_collider) colliders(g, sepset_len, sepset);
      neighbourhood_step();
    }
    std::memcpy(r_->pdag, pdag_.data(), sizeof(int) * p * p);
    if (opt_.perm_stream == KPC_PERM_R_STREAM && rstate_valid_)
      std::memcpy(r_->r_state_out, rstate_, sizeof(rstate_));
  }

 private:
  int at(int i, int j) const { return pdag_[(size_t)i * p_ + j]; }
  int& at(int i, int j) { return pdag_[(size_t)i * p_ + j]; }
  bool directed(int i, int j) const { return at(i, j) == 1 && at(j, i) == 0; }
  bool undirected(int i, int j) const { return at(i, j) == 1 && at(j, i) == 1; }
  bool adjacent(int i, int j) const { return at(i, j) + at(j, i) > 0; }

  // collider step: every unshielded x - y - z with y outside sepset(x, z)
  void colliders(const std::vector<int>& g, const int* slen, const int* sep) {
    const int p = p_;
    for (int y = 0; y < p; ++y)
      for (int x = 0; x < p; ++x) {
        if (g[(size_t)x * p + y] != 1) continue;
        for (int z = 0; z < p; ++z) {
          if (z == x || g[(size_t)y * p + z] != 1 || g[(size_t)x * p + z] != 0) continue;
          bool in_sep = false;
          for (int k = 0; k < 2 && !in_sep; ++k) {
            const int u = k ? z : x, v = k ? x : z;
            for (int s = 0; s < slen[(size_t)u * p + v]; ++s)
              in_sep = in_sep || sep[((size_t)u * p + v) * 64 + s] == y;
          }
          if (in_sep) continue;
          if (!opt_.solve_confl) {
            at(x, y) = at(z, y) = 1;
            at(y, x) = at(y, z) = 0;
          } else {
            if (at(x, y) == 1) at(y, x) = 0;
            else at(x, y) = at(y, x) = 2;
            if (at(z, y) == 1) at(y, z) = 0;
            else at(z, y) = at(y, z) = 2;
          }
        }
      }
  }

  // S a clique of the undirected skeleton and every parent of V adjacent to
  // every member of S (kpcalg's checkImmor)
  bool check_immor(int V, const std::vector<int>& S) const {
    const int k = (int)S.size();
    if (k > 1)
      for (int a = 0; a < k; ++a)
        for (int b = 0; b < k; ++b)
          if (a != b && !adjacent(S[a], S[b])) return false;
    int total = 0, np = 0;
    for (int i = 0; i < p_; ++i) {
      if (!directed(i, V)) continue;
      ++np;
      for (int s : S) total += adjacent(i, s);
    }
    return total >= k * np;
  }

  void rebuild_nbhd() {
    nb_.assign((size_t)p_, {});
    for (int i = 0; i < p_; ++i)
      for (int j = 0; j < p_; ++j)
        if (undirected(i, j)) nb_[i].push_back(j);
  }

  static bool next_subset(std::vector<int>& idx, int len) {
    const int k = (int)idx.size();
    int i = k - 1;
    while (i >= 0 && idx[i] == len - k + i) --i;
    if (i < 0) return false;
    ++idx[i];
    for (int j = i + 1; j < k; ++j) idx[j] = idx[j - 1] + 1;
    return true;
  }

  // the pairs of a predicate in which(arr.ind = TRUE) order
  template <typename F>
  std::vector<std::pair<int, int>> which_pairs(F f) const {
    std::vector<std::pair<int, int>> out;
    for (int j = 0; j < p_; ++j)
      for (int i = 0; i < p_; ++i)
        if (f(i, j)) out.emplace_back(i, j);
    return out;
  }

  void rule1() {
    const bool solve = opt_.solve_confl != 0;
    for (auto ab : which_pairs([this](int i, int j) { return directed(i, j); })) {
      const int a = ab.first, b = ab.second;
      std::vector<int> isC;
      for (int c = 0; c < p_; ++c)
        if (undirected(b, c) && at(a, c) == 0 && at(c, a) == 0) isC.push_back(c);
      for (int c : isC) {
        if (!solve || (at(b, c) == 1 && at(c, b) == 1)) {
          at(b, c) = 1;
          at(c, b) = 0;
        } else if (at(b, c) == 0 && at(c, b) == 1) {
          at(b, c) = at(c, b) = 2;
        }
      }
    }
  }

  void rule2() {
    const bool solve = opt_.solve_confl != 0;
    for (auto ab : which_pairs([this](int i, int j) { return undirected(i, j); })) {
      const int a = ab.first, b = ab.second;
      int chains = 0;
      for (int c = 0; c < p_; ++c) chains += directed(a, c) && directed(c, b);
      for (int k = 0; k < chains; ++k) {
        if (!solve || (at(a, b) == 1 && at(b, a) == 1)) {
          at(a, b) = 1;
          at(b, a) = 0;
        } else if (at(a, b) == 0 && at(b, a) == 1) {
          at(a, b) = at(b, a) = 2;
        }
      }
    }
  }

  void rule3() {
    const bool solve = opt_.solve_confl != 0;
    for (auto ab : which_pairs([this](int i, int j) { return undirected(i, j); })) {
      const int a = ab.first, b = ab.second;
      std::vector<int> c;
      for (int k = 0; k < p_; ++k)
        if (undirected(a, k) && directed(k, b)) c.push_back(k);
      bool done = false;
      for (size_t u = 0; u + 1 < c.size() && !done; ++u)
        for (size_t v = u + 1; v < c.size() && !done; ++v) {
          if (at(c[u], c[v]) != 0 || at(c[v], c[u]) != 0) continue;
          if (!solve || (at(a, b) == 1 && at(b, a) == 1)) {
            at(a, b) = 1;
            at(b, a) = 0;
            done = true;
          } else if (at(a, b) == 0 && at(b, a) == 1) {
            at(a, b) = at(b, a) = 2;
            done = true;
          }
        }
    }
  }

  // kpcalg's generalised transitive step
  void neighbourhood_step() {
    const int p = p_;
    rebuild_nbhd();
    std::vector<char> updt((size_t)p, 0);
    for (int s = 1;; ++s) {
      size_t mx = 0;
      for (const auto& v : nb_) mx = std::max(mx, v.size());
      if ((int)mx < s) break;
      for (int V = 0; V < p; ++V) {
        const int sz = (int)nb_[V].size();
        if (!(sz == s || (sz < s && updt[V]))) continue;
        int s2 = sz;
        while (s2 > 0) {
          const std::vector<int> vnb = nb_[V];  // the subsets come from this list
          const int len = (int)vnb.size();
          const int k = len == 1 ? 1 : s2;
          std::vector<int> idx((size_t)k);
          for (int i = 0; i < k; ++i) idx[i] = i;
          bool more = true;
          while (more) {
            std::vector<int> S((size_t)k);
            for (int i = 0; i < k; ++i) S[i] = vnb[idx[i]];
            bool to_update = true;
            if (check_immor(V, S)) {
              if (regr_v_on_ps(V, S) > 0) {
                to_update = false;
              } else {
                for (int W : S) {
                  int s3 = (int)nb_[W].size();
                  bool to_continue = true;
                  while (to_continue && s3 > 0) {
                    const std::vector<int> wnb = nb_[W];
                    const int wl = (int)wnb.size();
                    const int k3 = wl == 1 ? 1 : s3;
                    std::vector<int> j3((size_t)k3);
                    for (int i = 0; i < k3; ++i) j3[i] = i;
                    bool more3 = true;
                    while (more3) {
                      std::vector<int> S2((size_t)k3);
                      int hasV = 0;
                      for (int i = 0; i < k3; ++i) hasV += (S2[i] = wnb[j3[i]]) == V;
                      if (check_immor(W, S2) && hasV == 1 && regr_v_on_ps(W, S2) == 0) {
                        to_update = false;
                        to_continue = false;
                      }
                      more3 = wl > 1 && next_subset(j3, wl);
                    }
                    --s3;
                  }
                }
              }
              if (to_update) {
                const std::vector<int> cur = nb_[V];  // the neighbourhood as it stands now
                for (int v : S) {
                  at(V, v) = 0;
                  at(v, V) = 1;
                }
                for (int v : cur) {
                  bool inS = false;
                  for (int w : S) inS = inS || v == w;
                  if (inS) continue;
                  at(v, V) = 0;
                  at(V, v) = 1;
                }
                s2 = 0;
                for (;;) {
                  const std::vector<int> old = pdag_;
                  if (opt_.rules[0]) rule1();
                  if (opt_.rules[1]) rule2();
                  if (opt_.rules[2]) rule3();
                  if (old == pdag_) break;
                }
                rebuild_nbhd();
                updt[V] = 1;
              }
            }
            more = len > 1 && next_subset(idx, len);
          }
          --s2;
        }
      }
    }
  }

  // regrVonPS: the number of residual tests with p < alpha
  int regr_v_on_ps(int V, const std::vector<int>& S) {
    const int n = n_;
    std::vector<int> parents, cols(S);
    for (int i = 0; i < p_; ++i)
      if (directed(i, V)) parents.push_back(i);
    cols.insert(cols.end(), parents.begin(), parents.end());
    const int call = push_call(V, S, parents);
    // the residual of V on c(S, parents), GAM terms in that order
    std::vector<double> resid((size_t)n), info(32);
    const int nc = (int)cols.size(), off = 0;
    int rc = kpc_regrxons_mixed(data_, n, p_, 1, &V, &nc, &off, cols.data(), resid.data(),
                                info.data());
    if (rc) throw OrientError(rc, "orientation regrXonS failed", true);
    // one unconditional test of the residual against each raw column of S:
    // the matrix [resid | data[, S]] and tests (0, i) with S empty
    const int k = (int)S.size();
    std::vector<double> m((size_t)n * (1 + k));
    std::memcpy(m.data(), resid.data(), sizeof(double) * n);
    for (int i = 0; i < k; ++i)
      std::memcpy(m.data() + (size_t)(1 + i) * n, data_ + (size_t)S[i] * n, sizeof(double) * n);
    kpc_options o;
    kpc_default_options(&o);
    o.ci_method = opt_.ci_method;
    o.alpha = opt_.alpha;
    o.num_col = opt_.num_col;
    o.hsic_sig = opt_.hsic_sig;
    o.index = opt_.index;
    o.replicates = opt_.replicates;
    o.include_observed = opt_.include_observed;
    o.perm_stream = opt_.perm_stream;
    o.has_seed = 1;
    if (opt_.perm_stream == KPC_PERM_R_STREAM) {
      o.seed = opt_.seed;
      o.r_state = rstate_valid_ ? rstate_ : nullptr;
    } else {
      // counter-based indices: a per-call seed from (seed, method, V, sorted
      // c(S, parents)); the call's test i then keys (0, 1 + i) under it
      std::vector<int> cond(cols);
      std::sort(cond.begin(), cond.end());
      const uint64_t a = (uint64_t)V, b = (uint64_t)V;
      uint64_t h = splitmix(((uint64_t)opt_.seed ^ 0x6f7269656e74ULL) ^
                            (uint64_t)(opt_.ci_method + 1));
      h = splitmix(h ^ (a + 1));
      h = splitmix(h ^ (b + 1));
      h = splitmix(h ^ (uint64_t)(cond.size() + 1));
      for (int c : cond) h = splitmix(h ^ (uint64_t)(c + 1));
      o.seed = h;
    }
    std::vector<int> xs((size_t)k, 0), ys((size_t)k);
    for (int i = 0; i < k; ++i) ys[i] = 1 + i;
    std::vector<double> out((size_t)k * 4);
    int rs[625];
    rc = kpc_ci_batch_rs(m.data(), n, 1 + k, &o, k, xs.data(), ys.data(), 0, nullptr, out.data(),
                         rs);
    if (rc) throw OrientError(rc, "orientation CI test failed", true);
    if (opt_.perm_stream == KPC_PERM_R_STREAM &&
        (opt_.ci_method == KPC_DCC_PERM || opt_.ci_method == KPC_HSIC_PERM)) {
      std::memcpy(rstate_, rs, sizeof(rstate_));
      rstate_valid_ = true;
    }
    int cnt = 0;
    for (int i = 0; i < k; ++i) {
      double pv = out[(size_t)4 * i];
      if (std::isnan(pv)) pv = 1.0;
      out[(size_t)4 * i] = pv;
      push_test(call, S[i], &out[(size_t)4 * i]);
      cnt += pv < opt_.alpha;
    }
    r_->call_reject[call] = cnt;
    return cnt;
  }

  int push_call(int V, const std::vector<int>& S, const std::vector<int>& parents) {
    const int k = r_->n_calls++;
    if (k >= cap_calls_) {
      cap_calls_ = cap_calls_ ? 2 * cap_calls_ : 256;
      r_->call_v = (int*)std::realloc(r_->call_v, sizeof(int) * cap_calls_);
      r_->call_slen = (int*)std::realloc(r_->call_slen, sizeof(int) * cap_calls_);
      r_->call_S = (int*)std::realloc(r_->call_S, sizeof(int) * (size_t)cap_calls_ * 64);
      r_->call_plen = (int*)std::realloc(r_->call_plen, sizeof(int) * cap_calls_);
      r_->call_parents = (int*)std::realloc(r_->call_parents, sizeof(int) * (size_t)cap_calls_ * 64);
      r_->call_reject = (int*)std::realloc(r_->call_reject, sizeof(int) * cap_calls_);
    }
    r_->call_v[k] = V;
    r_->call_slen[k] = (int)S.size();
    r_->call_plen[k] = (int)parents.size();
    for (int i = 0; i < 64; ++i) {
      r_->call_S[(size_t)k * 64 + i] = i < (int)S.size() ? S[i] : 0;
      r_->call_parents[(size_t)k * 64 + i] = i < (int)parents.size() ? parents[i] : 0;
    }
    r_->call_reject[k] = 0;
    return k;
  }

  void push_test(int call, int other, const double* o4) {
    const int k = r_->n_tests++;
    if (k >= cap_tests_) {
      cap_tests_ = cap_tests_ ? 2 * cap_tests_ : 512;
      r_->test_call = (int*)std::realloc(r_->test_call, sizeof(int) * cap_tests_);
      r_->test_other = (int*)std::realloc(r_->test_other, sizeof(int) * cap_tests_);
      r_->test_out = (double*)std::realloc(r_->test_out, sizeof(double) * (size_t)cap_tests_ * 4);
    }
    r_->test_call[k] = call;
    r_->test_other[k] = other;
    for (int i = 0; i < 4; ++i) r_->test_out[(size_t)k * 4 + i] = o4[i];
  }

  const double* data_;
  int n_, p_;
  kpc_orient_options opt_;
  kpc_orient_result* r_;
  std::vector<int> pdag_;
  std::vector<std::vector<int>> nb_;
  int rstate_[625] = {};
  bool rstate_valid_ = false;
  int cap_calls_ = 0, cap_tests_ = 0;
};

}  // namespace

extern "C" {

void kpc_default_orient_options(kpc_orient_options* o) {
  std::memset(o, 0, sizeof(*o));
  o->ci_method = KPC_DCC_GAMMA;
  o->alpha = 0.1;
  o->num_col = 35;
  o->hsic_sig = 1.0;
  o->index = 1.0;
  o->replicates = 100;
  o->include_observed = 1;
  o->orient_collider = 1;
  o->rules[0] = o->rules[1] = o->rules[2] = 1;
}

int kpc_wanpdag_run(const double* data, int n, int p, const unsigned char* adjacency,
                    const int* sepset_len, const int* sepset, const kpc_orient_options* opt,
                    kpc_orient_result** out) {
  kpc_orient_result* r = nullptr;
  try {
    if (!data || !adjacency || !sepset_len || !sepset || !opt || !out)
      throw OrientError(-1, "null argument");
    if (p < 2 || p > 64) throw OrientError(-1, "p must be in [2, 64]");
    if (!(opt->alpha > 0.0 && opt->alpha < 1.0)) throw OrientError(-1, "alpha must be in (0, 1)");
    for (int i = 0; i < p; ++i)
      for (int j = 0; j < p; ++j)
        if ((adjacency[(size_t)i * p + j] != 0) != (adjacency[(size_t)j * p + i] != 0) ||
            (i == j && adjacency[(size_t)i * p + j]))
          throw OrientError(-1, "adjacency must be a symmetric skeleton with an empty diagonal");
    r = (kpc_orient_result*)std::calloc(1, sizeof(kpc_orient_result));
    r->p = p;
    r->pdag = (int*)std::malloc(sizeof(int) * (size_t)p * p);
    const auto t0 = std::chrono::steady_clock::now();
    Orienter(data, n, p, *opt, r).run(adjacency, sepset_len, sepset);
    r->wall_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    *out = r;
    return 0;
  } catch (const OrientError& e) {
    if (r) kpc_orient_result_free(r);
    if (!e.lib_error) kpc::set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    if (r) kpc_orient_result_free(r);
    kpc::set_last_error(e.what());
    return -100;
  }
}

void kpc_orient_result_free(kpc_orient_result* r) {
  if (!r) return;
  std::free(r->pdag);
  std::free(r->call_v);
  std::free(r->call_slen);
  std::free(r->call_S);
  std::free(r->call_plen);
  std::free(r->call_parents);
  std::free(r->call_reject);
  std::free(r->test_call);
  std::free(r->test_other);
  std::free(r->test_out);
  std::free(r);
}

}  // extern "C"
