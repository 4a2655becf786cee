/*
 * pcalg "stable" skeleton driven by kernelCItest, for the CPU oracle (test
 * infrastructure only).
 *
 * Follows fastkpc/src/skeleton_engine.cpp:96-227 (CPU twin of pcalg
 * skeleton(method="stable")) and fastkpc/src/skeleton_task_scheduler.cpp:
 * 82-121 (task order): per level take an adjacency snapshot; for every edge
 * x<y test x against subsets of adj(x)\{y} in lexicographic order, stop at the
 * first p >= alpha (delete, sepset = S), otherwise test y against subsets of
 * adj(y)\{x}; deletions apply at level end; pcalg's `done` flag ends the
 * search when no visited orientation had more than `level` neighbours.
 * kernelCItest semantics: kpcalg/R/kernelCItest.R:110-128 (raw columns at
 * |S| = 0, regrXonS residuals otherwise).
 */
#include "kpc_oracle.h"
#include "orc_internal.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Per-(target, S) component.  dcc.gamma: top-k eigenpairs (values, vecs),
 * dsum, moment.  hsic.*: centred inchol factor in vecs (rank columns),
 * dsum = off-diagonal kernel sum, moment = self moment.  dcc.perm: the
 * residual itself in vecs. */
typedef struct {
  uint64_t key_mask;
  int target;
  int used;
  double* values;
  double* vecs;
  double dsum, moment, margin;
  int rank;
} comp_entry;

typedef struct {
  comp_entry* slots;
  size_t cap;
  size_t count;
} comp_cache;

static uint64_t mix(uint64_t x) {
  x ^= x >> 33; x *= 0xff51afd7ed558ccdULL; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL; x ^= x >> 33;
  return x;
}

static comp_entry* cache_find(comp_cache* c, int target, uint64_t mask, int insert) {
  if (insert && (c->count + 1) * 2 > c->cap) {
    size_t ncap = c->cap ? c->cap * 2 : 1024;
    comp_entry* ns = (comp_entry*)calloc(ncap, sizeof(comp_entry));
    for (size_t i = 0; i < c->cap; ++i) {
      if (!c->slots[i].used) continue;
      size_t h = mix(c->slots[i].key_mask * 131ULL + (uint64_t)c->slots[i].target) & (ncap - 1);
      while (ns[h].used) h = (h + 1) & (ncap - 1);
      ns[h] = c->slots[i];
    }
    free(c->slots);
    c->slots = ns;
    c->cap = ncap;
  }
  if (!c->cap) return NULL;
  size_t h = mix(mask * 131ULL + (uint64_t)target) & (c->cap - 1);
  while (c->slots[h].used) {
    if (c->slots[h].target == target && c->slots[h].key_mask == mask) return &c->slots[h];
    h = (h + 1) & (c->cap - 1);
  }
  if (!insert) return NULL;
  c->slots[h].used = 1;
  c->slots[h].target = target;
  c->slots[h].key_mask = mask;
  c->count++;
  return &c->slots[h];
}

static void cache_free(comp_cache* c) {
  for (size_t i = 0; i < c->cap; ++i)
    if (c->slots[i].used) { free(c->slots[i].values); free(c->slots[i].vecs); }
  free(c->slots);
}

static uint64_t mask_of(const int* S, int s_len) {
  uint64_t m = 0;
  for (int i = 0; i < s_len; ++i) m |= 1ULL << S[i];
  return m;
}

/* residual of `target` on S (raw column when S is empty) */
static int residual_for(const double* data, int n, int target, const int* S, int s_len,
                        double* out, double* margin) {
  *margin = 1e300;
  if (s_len == 0) {
    memcpy(out, data + (size_t)target * n, sizeof(double) * n);
    return 0;
  }
  orc_magic_control ctl;
  orc_magic_default_control(&ctl);
  orc_magic_result res;
  int rc = orc_regrxons(data, n, n, target, S, s_len, &ctl, out, &res);
  *margin = res.margin;
  return rc;
}

typedef struct {
  double* values;
  double* vecs;
  double dsum, moment, margin;
  int rank;
  int rc;
} comp_val;

static comp_val compute_component(const double* data, int n, int target, const int* S,
                                  int s_len, const orc_skeleton_options* opt) {
  const int k = opt->num_col;
  comp_val v;
  memset(&v, 0, sizeof(v));
  v.values = (double*)malloc(sizeof(double) * k);
  v.vecs = (double*)malloc(sizeof(double) * (size_t)n * k);
  double* r = (double*)malloc(sizeof(double) * n);
  v.rc = residual_for(data, n, target, S, s_len, r, &v.margin);
  if (v.rc == 0) {
    if (opt->ci_method == ORC_DCC_GAMMA) {
      v.rc = orc_dcov_component(r, n, k, v.values, v.vecs, &v.dsum, &v.moment);
      v.rank = k;
    } else if (opt->ci_method == ORC_HSIC_GAMMA || opt->ci_method == ORC_HSIC_PERM) {
      double im = 1e300;
      v.rc = orc_hsic_component(r, n, opt->hsic_sig, k, v.vecs, &v.rank, &v.dsum, &v.moment,
                                &im);
      if (im < v.margin) v.margin = im;
    } else {
      memcpy(v.vecs, r, sizeof(double) * n);
      v.rank = 1;
    }
  }
  free(r);
  return v;
}

static int get_component(comp_cache* c, const double* data, int n, int target, const int* S,
                         int s_len, const orc_skeleton_options* opt, comp_entry** out) {
  uint64_t m = mask_of(S, s_len);
  comp_entry* e = cache_find(c, target, m, 0);
  if (!e) {
    comp_val v = compute_component(data, n, target, S, s_len, opt);
    if (v.rc != 0) { free(v.values); free(v.vecs); return v.rc; }
    e = cache_find(c, target, m, 1);
    e->values = v.values; e->vecs = v.vecs; e->dsum = v.dsum; e->moment = v.moment;
    e->margin = v.margin; e->rank = v.rank;
  }
  *out = e;
  return 0;
}

/* One CI test on cached components; out = {p, statistic, mean, variance,
 * margin}.  Permutation tests permute one variable: by default the one with
 * the larger index, with indices from the task seed of (x, y, S); with an R
 * stream (rng != NULL) the orientation's y, as kpcalg's kernelCItest passes
 * (x, y) to dcov.test / hsic.perm, with the test's table drawn from the
 * stream (fastkpc/src/full_cuda_ci_one_call.cpp:464-567). */
static int ci_pvalue(comp_cache* c, const double* data, int n, int x, int y, const int* S,
                     int s_len, const orc_skeleton_options* opt, orc_rng* rng, double out[5]) {
  if (opt->ci_method < ORC_DCC_GAMMA || opt->ci_method > ORC_HSIC_PERM) return -20;
  const int lo = rng ? x : (x < y ? x : y), hi = rng ? y : (x < y ? y : x);
  comp_entry *ex, *ey;
  int rc = get_component(c, data, n, lo, S, s_len, opt, &ex);
  if (rc) return rc;
  rc = get_component(c, data, n, hi, S, s_len, opt, &ey);
  if (rc) return rc;
  const unsigned long long seed =
      orc_task_seed(opt->perm_seed, opt->ci_method, x, y, S, s_len);
  int* table = NULL;
  if (rng && (opt->ci_method == ORC_DCC_PERM || opt->ci_method == ORC_HSIC_PERM)) {
    const int R = opt->replicates > 0 ? opt->replicates : 0;
    table = (int*)malloc(sizeof(int) * ((size_t)R * n + n));
    orc_r_perm_table(rng, opt->ci_method, n, R, table + (size_t)R * n, table);
  }
  double o[5];
  switch (opt->ci_method) {
    case ORC_DCC_GAMMA: {
      orc_dcov_pair_from_components(n, opt->num_col, ex->vecs, ex->values, ex->dsum, ex->moment,
                                    ey->vecs, ey->values, ey->dsum, ey->moment, o);
      o[4] = 1e300;
      break;
    }
    case ORC_HSIC_GAMMA:
      orc_hsic_gamma_from_factors(ex->vecs, ex->rank, ex->dsum, ex->moment, ey->vecs, ey->rank,
                                  ey->dsum, ey->moment, n, o);
      break;
    case ORC_HSIC_PERM:
      orc_hsic_perm_table(ex->vecs, ex->rank, ey->vecs, ey->rank, n, opt->replicates,
                          opt->include_observed, seed, table, o);
      break;
    default:
      rc = orc_dcc_perm_table(ex->vecs, ey->vecs, n, opt->index, opt->replicates,
                              opt->include_observed, seed, table, o);
      break;
  }
  free(table);
  if (rc) return rc;
  for (int i = 0; i < 4; ++i) out[i] = o[i];
  double mg = o[4];
  if (ex->margin < mg) mg = ex->margin;
  if (ey->margin < mg) mg = ey->margin;
  out[4] = mg;
  return 0;
}

int orc_ci_test(const double* data, int n, int p, int x, int y, const int* S, int s_len,
                const orc_skeleton_options* opt, double out[5]) {
  (void)p;
  comp_cache c = {0};
  int rc = ci_pvalue(&c, data, n, x, y, S, s_len, opt, NULL, out);
  cache_free(&c);
  return rc;
}

/* lexicographic k-subsets of nb[0..len-1] */
static int next_comb(int* idx, int k, int len) {
  int i = k - 1;
  while (i >= 0 && idx[i] == len - k + i) --i;
  if (i < 0) return 0;
  ++idx[i];
  for (int j = i + 1; j < k; ++j) idx[j] = idx[j - 1] + 1;
  return 1;
}

static void trace_push(orc_skeleton_result* r, int level, int x, int y, const int* S, int s_len,
                       double pv, int del, const double out[5]) {
  if (r->trace_len == r->trace_cap) {
    int nc = r->trace_cap ? r->trace_cap * 2 : 4096;
    r->trace_level = (int*)realloc(r->trace_level, sizeof(int) * nc);
    r->trace_x = (int*)realloc(r->trace_x, sizeof(int) * nc);
    r->trace_y = (int*)realloc(r->trace_y, sizeof(int) * nc);
    r->trace_slen = (int*)realloc(r->trace_slen, sizeof(int) * nc);
    r->trace_S = (int*)realloc(r->trace_S, sizeof(int) * (size_t)nc * 64);
    r->trace_p = (double*)realloc(r->trace_p, sizeof(double) * nc);
    r->trace_del = (unsigned char*)realloc(r->trace_del, nc);
    r->trace_margin = (double*)realloc(r->trace_margin, sizeof(double) * nc);
    r->trace_stat = (double*)realloc(r->trace_stat, sizeof(double) * nc);
    r->trace_mean = (double*)realloc(r->trace_mean, sizeof(double) * nc);
    r->trace_var = (double*)realloc(r->trace_var, sizeof(double) * nc);
    r->trace_cap = nc;
  }
  int t = r->trace_len++;
  r->trace_level[t] = level;
  r->trace_x[t] = x;
  r->trace_y[t] = y;
  r->trace_slen[t] = s_len;
  for (int i = 0; i < s_len; ++i) r->trace_S[(size_t)t * 64 + i] = S[i];
  r->trace_p[t] = pv;
  r->trace_del[t] = (unsigned char)del;
  r->trace_margin[t] = out[4];
  r->trace_stat[t] = out[1];
  r->trace_mean[t] = out[2];
  r->trace_var[t] = out[3];
}

/* Fill the component cache for every (target, S) the level could request. */
static int prefetch_level(comp_cache* c, const double* data, int n, int p, const unsigned char* snap,
                          int level, const orc_skeleton_options* opt) {
  int cap = 1024, cnt = 0;
  int* tg = (int*)malloc(sizeof(int) * cap);
  int* sl = (int*)malloc(sizeof(int) * (size_t)cap * 64);
  int nb[64], idx[64], S[64];
  for (int x = 0; x < p; ++x)
    for (int y = 0; y < p; ++y) {
      if (x == y || !snap[x * p + y]) continue;
      int len = 0;
      for (int i = 0; i < p; ++i) if (i != y && snap[i * p + x]) nb[len++] = i;
      if (len < level) continue;
      for (int i = 0; i < level; ++i) idx[i] = i;
      do {
        for (int i = 0; i < level; ++i) S[i] = nb[idx[i]];
        for (int side = 0; side < 2; ++side) {
          int t = side ? y : x;
          if (cache_find(c, t, mask_of(S, level), 0)) continue;
          int dup = 0;
          uint64_t m = mask_of(S, level);
          for (int q = cnt - 1; q >= 0 && q >= cnt - 8; --q)
            if (tg[q] == t && mask_of(sl + (size_t)q * 64, level) == m) dup = 1;
          if (dup) continue;
          if (cnt == cap) {
            cap *= 2;
            tg = (int*)realloc(tg, sizeof(int) * cap);
            sl = (int*)realloc(sl, sizeof(int) * (size_t)cap * 64);
          }
          tg[cnt] = t;
          memcpy(sl + (size_t)cnt * 64, S, sizeof(int) * level);
          ++cnt;
          /* reserve the key so later duplicates are skipped */
          comp_entry* e = cache_find(c, t, m, 1);
          e->values = NULL;
          e->vecs = NULL;
        }
      } while (level > 0 && next_comb(idx, level, len));
    }
  int rc = 0;
  comp_val* vals = (comp_val*)calloc(cnt ? cnt : 1, sizeof(comp_val));
#pragma omp parallel for schedule(dynamic, 1) num_threads(opt->threads > 0 ? opt->threads : 1)
  for (int i = 0; i < cnt; ++i)
    vals[i] = compute_component(data, n, tg[i], sl + (size_t)i * 64, level, opt);
  for (int i = 0; i < cnt; ++i) {
    comp_entry* e = cache_find(c, tg[i], mask_of(sl + (size_t)i * 64, level), 0);
    if (vals[i].rc != 0) { rc = vals[i].rc; free(vals[i].values); free(vals[i].vecs); continue; }
    e->values = vals[i].values;
    e->vecs = vals[i].vecs;
    e->dsum = vals[i].dsum;
    e->moment = vals[i].moment;
    e->margin = vals[i].margin;
    e->rank = vals[i].rank;
  }
  free(vals); free(tg); free(sl);
  return rc;
}

/* One orientation (x, y) of an edge a < b at a level: its neighbours
 * adj(x) \ {y} in the level's snapshot and its tasks' span in the canonical
 * plan (edges a < b row-major, the (a, b) orientation's conditioning sets
 * before the (b, a) orientation's, each in lexicographic order). */
typedef struct {
  int a, b, x, y, len, first, count;
  int nb[64];
} orient_t;

typedef struct {
  int task;  /* canonical plan position */
  int x, y, s_len, del;
  int S[64];
  double out[5];
} done_test;

static long long n_choose(int n, int k) {
  if (k < 0 || k > n) return 0;
  long long r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

static int cmp_pcalg(const void* l, const void* r) {
  const orient_t* u = *(const orient_t* const*)l;
  const orient_t* v = *(const orient_t* const*)r;
  if (u->x != v->x) return u->x < v->x ? -1 : 1;
  return u->y < v->y ? -1 : (u->y > v->y);
}

static int cmp_task(const void* l, const void* r) {
  const done_test* u = (const done_test*)l;
  const done_test* v = (const done_test*)r;
  return u->task < v->task ? -1 : (u->task > v->task);
}

int orc_skeleton(const double* data, int n, int p, const orc_skeleton_options* opt,
                 orc_skeleton_result* r) {
  memset(r, 0, sizeof(*r));
  if (p < 2 || p > 64) return -30;
  r->p = p;
  r->adjacency = (unsigned char*)malloc((size_t)p * p);
  r->pmax = (double*)malloc(sizeof(double) * (size_t)p * p);
  r->sepset_len = (int*)malloc(sizeof(int) * (size_t)p * p);
  r->sepset = (int*)calloc((size_t)p * p * 64, sizeof(int));
  for (int i = 0; i < p * p; ++i) { r->adjacency[i] = 1; r->pmax[i] = -INFINITY; r->sepset_len[i] = -1; }
  for (int i = 0; i < p; ++i) { r->adjacency[i * p + i] = 0; r->pmax[i * p + i] = 1.0; }
  unsigned char* snap = (unsigned char*)malloc((size_t)p * p);
  unsigned char* del = (unsigned char*)malloc((size_t)p * p);
  orient_t* ors = (orient_t*)malloc(sizeof(orient_t) * (size_t)p * (p - 1));
  orient_t** order = (orient_t**)malloc(sizeof(orient_t*) * (size_t)p * (p - 1));
  orc_rng rng;
  orc_rng* rp = NULL;
  if (opt->r_stream) {
    orc_rng_set_seed(&rng, (unsigned int)opt->perm_seed);
    rp = &rng;
  }
  comp_cache cache = {0};
  int idx[64], S[64];
  int done = 0, ord = 0, rc = 0;
  while (!done && ord <= opt->m_max) {
    int any = 0;
    for (int i = 0; i < p * p; ++i) any |= r->adjacency[i];
    if (!any) break;
    done = 1;
    memcpy(snap, r->adjacency, (size_t)p * p);
    memset(del, 0, (size_t)p * p);
    if (opt->threads > 0) {
      rc = prefetch_level(&cache, data, n, p, snap, ord, opt);
      if (rc) break;
    }
    /* the level plan's orientations, canonical order */
    int no = 0;
    long long planned = 0;
    for (int a = 0; a < p - 1; ++a)
      for (int b = a + 1; b < p; ++b) {
        if (!snap[a * p + b]) continue;
        for (int side = 0; side < 2; ++side) {
          orient_t* o = &ors[no++];
          o->a = a; o->b = b;
          o->x = side ? b : a;
          o->y = side ? a : b;
          o->len = 0;
          for (int i = 0; i < p; ++i) if (i != o->y && snap[i * p + o->x]) o->nb[o->len++] = i;
          o->first = (int)planned;
          o->count = (int)n_choose(o->len, ord);
          planned += o->count;
        }
      }
    /* execution order: canonical, or pcalg's (x, then y) for the R stream --
     * the order kpcalg runs the tests in and so draws their permutations */
    for (int i = 0; i < no; ++i) order[i] = &ors[i];
    if (rp) qsort(order, no, sizeof(orient_t*), cmp_pcalg);
    done_test* res = NULL;
    int nres = 0, cres = 0;
    for (int oi = 0; oi < no && !rc; ++oi) {
      const orient_t* o = order[oi];
      if (del[o->a * p + o->b]) continue;  /* deleted by its other orientation */
      if (o->len < ord) continue;
      if (o->len > ord) done = 0;
      for (int i = 0; i < ord; ++i) idx[i] = i;
      int k = 0;
      do {
        for (int i = 0; i < ord; ++i) S[i] = o->nb[idx[i]];
        if (nres == cres) {
          cres = cres ? 2 * cres : 1024;
          res = (done_test*)realloc(res, sizeof(done_test) * cres);
        }
        done_test* d = &res[nres++];
        rc = ci_pvalue(&cache, data, n, o->x, o->y, S, ord, opt, rp, d->out);
        if (rc) break;
        double pv = d->out[0];
        if (isnan(pv)) pv = 1.0;
        d->out[0] = pv;
        d->task = o->first + k++;
        d->x = o->x; d->y = o->y; d->s_len = ord;
        memcpy(d->S, S, sizeof(int) * ord);
        if (r->pmax[o->x * p + o->y] < pv) r->pmax[o->x * p + o->y] = pv;
        d->del = pv >= opt->alpha;
        if (d->del) {
          del[o->a * p + o->b] = del[o->b * p + o->a] = 1;
          r->sepset_len[o->a * p + o->b] = r->sepset_len[o->b * p + o->a] = ord;
          for (int i = 0; i < ord; ++i) {
            r->sepset[((size_t)o->a * p + o->b) * 64 + i] = S[i];
            r->sepset[((size_t)o->b * p + o->a) * 64 + i] = S[i];
          }
          break;
        }
      } while (ord > 0 && next_comb(idx, ord, o->len));
    }
    /* trace rows in canonical plan order */
    qsort(res, nres, sizeof(done_test), cmp_task);
    for (int i = 0; i < nres && !rc; ++i)
      trace_push(r, ord, res[i].x, res[i].y, res[i].S, res[i].s_len, res[i].out[0], res[i].del,
                 res[i].out);
    free(res);
    for (int i = 0; i < p * p; ++i) if (del[i]) r->adjacency[i] = 0;
    r->n_edgetests[ord] = nres;
    ++ord;
  }
  r->levels = ord;
  for (int i = 0; i < p - 1; ++i)
    for (int j = i + 1; j < p; ++j) {
      double m = r->pmax[i * p + j] > r->pmax[j * p + i] ? r->pmax[i * p + j] : r->pmax[j * p + i];
      r->pmax[i * p + j] = r->pmax[j * p + i] = m;
    }
  if (rp) r->rng_out = rng;
  cache_free(&cache);
  free(snap);
  free(del);
  free(ors);
  free(order);
  return rc;
}

void orc_skeleton_result_free(orc_skeleton_result* r) {
  free(r->adjacency); free(r->pmax); free(r->sepset_len); free(r->sepset);
  free(r->trace_level); free(r->trace_x); free(r->trace_y); free(r->trace_slen);
  free(r->trace_S); free(r->trace_p); free(r->trace_del); free(r->trace_margin);
  free(r->trace_stat); free(r->trace_mean); free(r->trace_var);
  memset(r, 0, sizeof(*r));
}
