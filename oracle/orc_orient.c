/*
 * kpcalg's WAN-PDAG orientation (generalised transitive step) for the CPU
 * oracle (test infrastructure only).
 *
 * Restates kpcalg/R/udag2wanpdag.R:104-436 (collider step, rules 1-3 with
 * solve.confl, checkImmor, the neighbourhood loop) and kpcalg/R/regrvonps.R:
 * 17-39 (regress V on c(S, parents(V)) in that column order, then one
 * unconditional CI test of the residual against each raw column of S; the
 * call returns how many p-values fall below alpha), as driven by fastkpc's
 * kpcalg-authority wrapper (fastkpc/R/fast_kpc.R:263-445).  R's iteration
 * order is kept exactly: which(M, arr.ind = TRUE) walks M column by column,
 * combn() enumerates subsets in lexicographic order of positions, the rule
 * passes iterate over index lists computed once per pass, and the loops of
 * the neighbourhood step do not stop early (a rejected orientation still
 * runs the remaining regressions of its loop).  Matrices are p x p row-major,
 * pdag[i * p + j] = R's pdag[i + 1, j + 1].
 */
#include "kpc_oracle.h"
#include "orc_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  const double* data;
  int n, p;
  const orc_orient_options* opt;
  orc_rng* rng;
  orc_orient_result* res;
  double* work;  /* 2 n: residual and scratch */
} orient_ctx;

static void push_call(orc_orient_result* r, int V, const int* S, int s_len, const int* par,
                      int n_par) {
  if (r->n_calls == r->call_cap) {
    int c = r->call_cap ? 2 * r->call_cap : 256;
    r->call_v = (int*)realloc(r->call_v, sizeof(int) * c);
    r->call_slen = (int*)realloc(r->call_slen, sizeof(int) * c);
    r->call_S = (int*)realloc(r->call_S, sizeof(int) * (size_t)c * 64);
    r->call_plen = (int*)realloc(r->call_plen, sizeof(int) * c);
    r->call_parents = (int*)realloc(r->call_parents, sizeof(int) * (size_t)c * 64);
    r->call_reject = (int*)realloc(r->call_reject, sizeof(int) * c);
    r->call_cap = c;
  }
  int k = r->n_calls++;
  r->call_v[k] = V;
  r->call_slen[k] = s_len;
  r->call_plen[k] = n_par;
  for (int i = 0; i < 64; ++i) {
    r->call_S[(size_t)k * 64 + i] = i < s_len ? S[i] : 0;
    r->call_parents[(size_t)k * 64 + i] = i < n_par ? par[i] : 0;
  }
  r->call_reject[k] = 0;
}

static void push_test(orc_orient_result* r, int call, int other, const double out[5]) {
  if (r->n_tests == r->test_cap) {
    int c = r->test_cap ? 2 * r->test_cap : 512;
    r->test_call = (int*)realloc(r->test_call, sizeof(int) * c);
    r->test_other = (int*)realloc(r->test_other, sizeof(int) * c);
    r->test_out = (double*)realloc(r->test_out, sizeof(double) * (size_t)c * 4);
    r->test_cap = c;
  }
  int k = r->n_tests++;
  r->test_call[k] = call;
  r->test_other[k] = other;
  for (int i = 0; i < 4; ++i) r->test_out[(size_t)k * 4 + i] = out[i];
}

/* one unconditional test of x against y (kernelCItest with S = NULL):
 * out = {p, statistic, mean, variance, margin} */
static int pair_test(orient_ctx* c, const double* x, const double* y, unsigned long long seed,
                     double out[5]) {
  const orc_orient_options* o = c->opt;
  const int n = c->n;
  int rc = 0;
  if (o->ci_method == ORC_DCC_GAMMA) {
    rc = orc_dcov_gamma(x, y, n, o->num_col, out);
    out[4] = 1e300;
    return rc;
  }
  int* table = NULL;
  if (o->ci_method == ORC_DCC_PERM || o->ci_method == ORC_HSIC_PERM) {
    if (o->r_stream) {
      table = (int*)malloc(sizeof(int) * ((size_t)o->replicates * n + n));
      orc_r_perm_table(c->rng, o->ci_method, n, o->replicates, table + (size_t)o->replicates * n,
                       table);
    }
  }
  if (o->ci_method == ORC_DCC_PERM) {
    rc = orc_dcc_perm_table(x, y, n, o->index, o->replicates, o->include_observed, seed, table,
                            out);
  } else {
    const int k = o->num_col;
    double* gx = (double*)malloc(sizeof(double) * (size_t)n * k);
    double* gy = (double*)malloc(sizeof(double) * (size_t)n * k);
    int rx = 0, ry = 0;
    double ox, mx, oy, my, mgx = 1e300, mgy = 1e300;
    rc = orc_hsic_component(x, n, o->hsic_sig, k, gx, &rx, &ox, &mx, &mgx);
    if (!rc) rc = orc_hsic_component(y, n, o->hsic_sig, k, gy, &ry, &oy, &my, &mgy);
    if (!rc) {
      if (o->ci_method == ORC_HSIC_GAMMA)
        orc_hsic_gamma_from_factors(gx, rx, ox, mx, gy, ry, oy, my, n, out);
      else
        orc_hsic_perm_table(gx, rx, gy, ry, n, o->replicates, o->include_observed, seed, table,
                            out);
      if (mgx < out[4]) out[4] = mgx;
      if (mgy < out[4]) out[4] = mgy;
    }
    free(gx);
    free(gy);
  }
  free(table);
  return rc;
}

/* the seed of the counter-based indices of residual test i of a regrVonPS
 * call on V with residual conditioning set cond (sorted c(S, parents)): a
 * per-call seed from (seed, method, V, cond), then the test keyed as (0,
 * 1 + i) of the matrix [residual | data[, S]] */
unsigned long long orc_orient_seed(unsigned long long user_seed, int method, int V, int i,
                                   const int* cond, int c_len) {
  const unsigned long long call =
      orc_task_seed(user_seed ^ 0x6f7269656e74ULL, method, V, V, cond, c_len);
  return orc_task_seed(call, method, 0, 1 + i, NULL, 0);
}

static int cmp_int(const void* a, const void* b) {
  const int u = *(const int*)a, v = *(const int*)b;
  return u < v ? -1 : (u > v);
}

/* regrVonPS(G = pdag, V, S): the number of residual tests with p < alpha */
static int regr_v_on_ps(orient_ctx* c, const int* pdag, int V, const int* S, int s_len,
                        int* reject) {
  const int p = c->p, n = c->n;
  int par[64], np = 0, cols[128], cond[128];
  for (int i = 0; i < p; ++i)
    if (pdag[i * p + V] == 1 && pdag[V * p + i] == 0) par[np++] = i;
  int nc = 0;
  for (int i = 0; i < s_len; ++i) cols[nc++] = S[i];
  for (int i = 0; i < np; ++i) cols[nc++] = par[i];
  push_call(c->res, V, S, s_len, par, np);
  const int call = c->res->n_calls - 1;
  orc_magic_control ctl;
  orc_magic_default_control(&ctl);
  orc_magic_result mr;
  int rc = orc_regrxons(c->data, n, n, V, cols, nc, &ctl, c->work, &mr);
  if (rc) return rc;
  memcpy(cond, cols, sizeof(int) * nc);
  qsort(cond, nc, sizeof(int), cmp_int);
  int cnt = 0;
  for (int i = 0; i < s_len; ++i) {
    double out[5];
    const unsigned long long seed =
        orc_orient_seed(c->opt->perm_seed, c->opt->ci_method, V, i, cond, nc);
    rc = pair_test(c, c->work, c->data + (size_t)S[i] * n, seed, out);
    if (rc) return rc;
    double pv = out[0];
    if (isnan(pv)) pv = 1.0;
    out[0] = pv;
    push_test(c->res, call, S[i], out);
    if (pv < c->opt->alpha) ++cnt;
  }
  c->res->call_reject[call] = cnt;
  *reject = cnt;
  return 0;
}

/* checkImmor(pdag, V, S): S is a clique of the undirected skeleton of pdag
 * and every parent of V is adjacent to every member of S */
static int check_immor(const int* pdag, int p, int V, const int* S, int s_len) {
#define UD(i, j) ((pdag[(i) * p + (j)] + pdag[(j) * p + (i)]) > 0)
  if (s_len > 1)
    for (int a = 0; a < s_len; ++a)
      for (int b = 0; b < s_len; ++b)
        if (a != b && !UD(S[a], S[b])) return 0;
  int total = 0, np = 0;
  for (int i = 0; i < p; ++i) {
    if (!(pdag[i * p + V] == 1 && pdag[V * p + i] == 0)) continue;
    ++np;
    for (int k = 0; k < s_len; ++k) total += UD(i, S[k]);
  }
  return total >= s_len * np;
#undef UD
}

static int next_subset(int* idx, int k, int len) {
  int i = k - 1;
  while (i >= 0 && idx[i] == len - k + i) --i;
  if (i < 0) return 0;
  ++idx[i];
  for (int j = i + 1; j < k; ++j) idx[j] = idx[j - 1] + 1;
  return 1;
}

/* which(M, arr.ind = TRUE) of a predicate on (i, j), column by column */
typedef int (*pair_pred)(const int* pdag, int p, int i, int j);
static int pred_directed(const int* g, int p, int i, int j) { return g[i * p + j] == 1 && g[j * p + i] == 0; }
static int pred_undirected(const int* g, int p, int i, int j) { return g[i * p + j] == 1 && g[j * p + i] == 1; }

static int which_pairs(const int* pdag, int p, pair_pred f, int* ia, int* ib) {
  int k = 0;
  for (int j = 0; j < p; ++j)
    for (int i = 0; i < p; ++i)
      if (f(pdag, p, i, j)) { ia[k] = i; ib[k] = j; ++k; }
  return k;
}

static void rule1(int* pdag, int p, int solve, int* ia, int* ib, int* isC) {
  const int m = which_pairs(pdag, p, pred_directed, ia, ib);
  for (int t = 0; t < m; ++t) {
    const int a = ia[t], b = ib[t];
    int nc = 0;
    for (int c = 0; c < p; ++c)
      if (pdag[b * p + c] == 1 && pdag[c * p + b] == 1 && pdag[a * p + c] == 0 &&
          pdag[c * p + a] == 0)
        isC[nc++] = c;
    for (int k = 0; k < nc; ++k) {
      const int c = isC[k];
      if (!solve || (pdag[b * p + c] == 1 && pdag[c * p + b] == 1)) {
        pdag[b * p + c] = 1;
        pdag[c * p + b] = 0;
      } else if (pdag[b * p + c] == 0 && pdag[c * p + b] == 1) {
        pdag[b * p + c] = 2;
        pdag[c * p + b] = 2;
      }
    }
  }
}

static void rule2(int* pdag, int p, int solve, int* ia, int* ib, int* isC) {
  const int m = which_pairs(pdag, p, pred_undirected, ia, ib);
  for (int t = 0; t < m; ++t) {
    const int a = ia[t], b = ib[t];
    int nc = 0;
    for (int c = 0; c < p; ++c)
      if (pdag[a * p + c] == 1 && pdag[c * p + a] == 0 && pdag[c * p + b] == 1 &&
          pdag[b * p + c] == 0)
        isC[nc++] = c;
    for (int k = 0; k < nc; ++k) {
      if (!solve || (pdag[a * p + b] == 1 && pdag[b * p + a] == 1)) {
        pdag[a * p + b] = 1;
        pdag[b * p + a] = 0;
      } else if (pdag[a * p + b] == 0 && pdag[b * p + a] == 1) {
        pdag[a * p + b] = 2;
        pdag[b * p + a] = 2;
      }
    }
  }
}

static void rule3(int* pdag, int p, int solve, int* ia, int* ib, int* isC) {
  const int m = which_pairs(pdag, p, pred_undirected, ia, ib);
  for (int t = 0; t < m; ++t) {
    const int a = ia[t], b = ib[t];
    int nc = 0;
    for (int c = 0; c < p; ++c)
      if (pdag[a * p + c] == 1 && pdag[c * p + a] == 1 && pdag[c * p + b] == 1 &&
          pdag[b * p + c] == 0)
        isC[nc++] = c;
    if (nc < 2) continue;
    int done = 0;
    for (int u = 0; u < nc - 1 && !done; ++u)
      for (int v = u + 1; v < nc && !done; ++v) {
        const int c1 = isC[u], c2 = isC[v];
        if (pdag[c1 * p + c2] == 0 && pdag[c2 * p + c1] == 0) {
          if (!solve || (pdag[a * p + b] == 1 && pdag[b * p + a] == 1)) {
            pdag[a * p + b] = 1;
            pdag[b * p + a] = 0;
            done = 1;
          } else if (pdag[a * p + b] == 0 && pdag[b * p + a] == 1) {
            pdag[a * p + b] = pdag[b * p + a] = 2;
            done = 1;
          }
        }
      }
  }
}

/* undirected neighbourhoods: nb[i] ascending, size[i] */
static void undirected_nbhd(const int* pdag, int p, int* nb, int* size) {
  for (int i = 0; i < p; ++i) {
    size[i] = 0;
    for (int j = 0; j < p; ++j)
      if (pdag[i * p + j] == 1 && pdag[j * p + i] == 1) nb[i * 64 + size[i]++] = j;
  }
}

int orc_wanpdag(const double* data, int n, int p, const unsigned char* adjacency,
                const int* sepset_len, const int* sepset, const orc_orient_options* opt,
                orc_orient_result* res) {
  memset(res, 0, sizeof(*res));
  if (p < 2 || p > 64) return -30;
  orient_ctx c;
  c.data = data;
  c.n = n;
  c.p = p;
  c.opt = opt;
  c.res = res;
  orc_rng rng;
  if (opt->r_stream) {
    if (opt->r_state_in) rng = *opt->r_state_in;
    else orc_rng_set_seed(&rng, (unsigned int)opt->perm_seed);
  }
  c.rng = &rng;
  c.work = (double*)malloc(sizeof(double) * 2 * (size_t)n);
  int* g = (int*)malloc(sizeof(int) * p * p);
  int* pdag = (int*)malloc(sizeof(int) * p * p);
  int* old = (int*)malloc(sizeof(int) * p * p);
  int* ia = (int*)malloc(sizeof(int) * p * p);
  int* ib = (int*)malloc(sizeof(int) * p * p);
  int* isC = (int*)malloc(sizeof(int) * p);
  int* nb = (int*)malloc(sizeof(int) * p * 64);
  int* size = (int*)malloc(sizeof(int) * p);
  int* updt = (int*)calloc(p, sizeof(int));
  int edges = 0;
  for (int i = 0; i < p * p; ++i) {
    g[i] = adjacency[i] ? 1 : 0;
    edges += g[i];
  }
  memcpy(pdag, g, sizeof(int) * p * p);
  int rc = 0;
  if (edges == 0) goto finish;
  if (opt->orient_collider) {
    const int m = which_pairs(g, p, pred_undirected, ia, ib);
    for (int t = 0; t < m; ++t) {
      const int x = ia[t], y = ib[t];
      for (int z = 0; z < p; ++z) {
        if (z == x || g[y * p + z] != 1) continue;
        if (g[x * p + z] != 0) continue;
        int in_sep = 0;
        for (int k = 0; k < 2 && !in_sep; ++k) {
          const int u = k ? z : x, v = k ? x : z;
          const int len = sepset_len[u * p + v];
          for (int s = 0; s < len; ++s) in_sep |= sepset[((size_t)u * p + v) * 64 + s] == y;
        }
        if (in_sep) continue;
        if (!opt->solve_confl) {
          pdag[x * p + y] = pdag[z * p + y] = 1;
          pdag[y * p + x] = pdag[y * p + z] = 0;
        } else {
          if (pdag[x * p + y] == 1) pdag[y * p + x] = 0;
          else pdag[x * p + y] = pdag[y * p + x] = 2;
          if (pdag[z * p + y] == 1) pdag[y * p + z] = 0;
          else pdag[z * p + y] = pdag[y * p + z] = 2;
        }
      }
    }
  }
  undirected_nbhd(pdag, p, nb, size);
  for (int s = 1;; ++s) {
    int mx = 0;
    for (int i = 0; i < p; ++i) if (size[i] > mx) mx = size[i];
    if (mx < s) break;
    for (int V = 0; V < p && !rc; ++V) {
      int s2 = s;
      if (!(size[V] == s2 || (size[V] < s2 && updt[V]))) continue;
      s2 = size[V];
      while (s2 > 0 && !rc) {
        /* the subsets of V's neighbourhood as it stands now */
        int vnb[64], vlen = size[V];
        memcpy(vnb, nb + V * 64, sizeof(int) * vlen);
        const int k = vlen == 1 ? 1 : s2;
        int idx[64];
        for (int i = 0; i < k; ++i) idx[i] = i;
        int more = 1;
        while (more && !rc) {
          int S[64];
          for (int i = 0; i < k; ++i) S[i] = vnb[idx[i]];
          int to_update = 1;
          if (check_immor(pdag, p, V, S, k)) {
            int rej = 0;
            rc = regr_v_on_ps(&c, pdag, V, S, k, &rej);
            if (rc) break;
            if (rej > 0) {
              to_update = 0;
            } else {
              for (int kk = 0; kk < k && !rc; ++kk) {
                const int W = S[kk];
                int s3 = size[W];
                int to_continue = 1;
                while (to_continue && s3 > 0 && !rc) {
                  int wnb[64], wlen = size[W];
                  memcpy(wnb, nb + W * 64, sizeof(int) * wlen);
                  const int k3 = wlen == 1 ? 1 : s3;
                  int j3[64];
                  for (int i = 0; i < k3; ++i) j3[i] = i;
                  int more3 = 1;
                  while (more3 && !rc) {
                    int S2[64], hasV = 0;
                    for (int i = 0; i < k3; ++i) {
                      S2[i] = wnb[j3[i]];
                      hasV += S2[i] == V;
                    }
                    if (check_immor(pdag, p, W, S2, k3) && hasV == 1) {
                      int rej2 = 0;
                      rc = regr_v_on_ps(&c, pdag, W, S2, k3, &rej2);
                      if (!rc && rej2 == 0) {
                        to_update = 0;
                        to_continue = 0;
                      }
                    }
                    more3 = wlen > 1 && next_subset(j3, k3, wlen);
                  }
                  --s3;
                }
              }
            }
            if (to_update && !rc) {
              /* S -> V, and V -> the rest of V's current neighbourhood (the
               * list as it stands now, rebuilt by any earlier acceptance) */
              int cur[64], clen = size[V];
              memcpy(cur, nb + V * 64, sizeof(int) * clen);
              for (int i = 0; i < k; ++i) {
                pdag[V * p + S[i]] = 0;
                pdag[S[i] * p + V] = 1;
              }
              for (int j = 0; j < clen; ++j) {
                int inS = 0;
                for (int i = 0; i < k; ++i) inS |= cur[j] == S[i];
                if (inS) continue;
                pdag[cur[j] * p + V] = 0;
                pdag[V * p + cur[j]] = 1;
              }
              s2 = 0;
              for (;;) {
                memcpy(old, pdag, sizeof(int) * p * p);
                if (opt->rules[0]) rule1(pdag, p, opt->solve_confl, ia, ib, isC);
                if (opt->rules[1]) rule2(pdag, p, opt->solve_confl, ia, ib, isC);
                if (opt->rules[2]) rule3(pdag, p, opt->solve_confl, ia, ib, isC);
                if (!memcmp(old, pdag, sizeof(int) * p * p)) break;
              }
              undirected_nbhd(pdag, p, nb, size);
              updt[V] = 1;
            }
          }
          more = vlen > 1 && next_subset(idx, k, vlen);
        }
        --s2;
      }
    }
    if (rc) break;
  }
finish:
  res->pdag = pdag;
  if (opt->r_stream) res->rng_out = rng;
  free(c.work); free(g); free(old); free(ia); free(ib); free(isC); free(nb); free(size);
  free(updt);
  return rc;
}

void orc_orient_result_free(orc_orient_result* r) {
  free(r->pdag);
  free(r->call_v); free(r->call_slen); free(r->call_S); free(r->call_plen);
  free(r->call_parents); free(r->call_reject);
  free(r->test_call); free(r->test_other); free(r->test_out);
  memset(r, 0, sizeof(*r));
}
