/*
 * kpcmi — MI355X (gfx950) kernel-PC conditional-independence path, C ABI.
 *
 * Drop-in boundary for the reference's R-facing native entry points:
 *   kpc_skeleton_run    replaces  C_full_cuda_ci_one_call_skeleton[_method]
 *                                 (fastkpc/src/r_api_cuda.cpp:9183-9249) and the
 *                                 C++ fastkpc::full_cuda_ci_one_call_skeleton_method
 *                                 (fastkpc/src/full_cuda_ci_one_call.hpp:19-36)
 *   kpc_ci_batch        replaces  one same-S cohort of execute_group
 *                                 (fastkpc/src/full_cuda_ci_one_call.cpp:4197-5036);
 *                                 per-test semantics of kpcalg kernelCItest
 *                                 (kpcalg/R/kernelCItest.R:68-129)
 *   kpc_regrxons_batch  replaces  the GCV optimiser + fixed-sp residual solve
 *                                 (mgcv_single_penalty_gcv.hpp / mgcv_multi_penalty_gcv_capacity.hpp /
 *                                  mgcv_fixed_sp_runtime.hpp:70-75); semantics of
 *                                 kpcalg regrXonS (kpcalg/R/regrXonS.R:33-61)
 *   kpc_gcv_grad_hess   exposes   the GCV score gradient/Hessian in log(sp)
 *                                 (fastkpc/src/mgcv_multi_penalty_cpp.cpp:819-990)
 *   kpc_permutation_table replaces full_cuda_ci_method_seeded_permutation_table
 *                                 (fastkpc/src/cuda/full_cuda_ci_method_batch.cu:1681-1717)
 *   kpc_task_seed       replaces  task_permutation_seed
 *                                 (fastkpc/src/full_cuda_ci_one_call.cpp:296-325)
 *   kpc_last_error      replaces  the std::runtime_error -> R error mapping
 *                                 (BEGIN_RCPP/END_RCPP in r_api_cuda.cpp)
 *
 * Conventions: every entry returns 0 on success and a negative status on
 * failure (then kpc_last_error() describes it); inputs are host pointers to
 * column-major n x p doubles unless the name says _device; outputs are caller
 * allocated except kpc_result, which kpc_result_free releases.  Calls from
 * several host threads may run at the same time on one GPU: each call has its
 * own streams and device arenas (a thread reuses only the arenas it released
 * itself), and the process-wide result / target-state caches are locked.
 */
#ifndef KPCMI_H
#define KPCMI_H

#ifdef __cplusplus
extern "C" {
#endif

enum kpc_ci_method { KPC_DCC_GAMMA = 0, KPC_HSIC_GAMMA = 1, KPC_DCC_PERM = 2, KPC_HSIC_PERM = 3 };
enum kpc_exec_mode { KPC_BATCHED = 0, KPC_PLAIN = 1 };
/* where the permutation methods take their indices from */
enum kpc_perm_stream {
  KPC_PERM_COUNTER = 0,  /* Philox4x32-10 keyed by the task seed (any order, on the device) */
  KPC_PERM_R_STREAM = 1  /* R's global Mersenne-Twister stream, drawn test by test in
                            pcalg's execution order as kpcalg does (host generated) */
};

typedef struct kpc_options {
  int ci_method;             /* kpc_ci_method */
  double alpha;              /* deletion threshold: p >= alpha deletes */
  int max_conditioning_size; /* m.max; negative means Inf (resolved to p - 2) */
  double index;              /* dcov index; outside [0, 2] reported as 1 */
  int num_col;               /* truncation rank, 0 < num_col < n */
  double hsic_sig;           /* rbfdot sigma = 1 / hsic_sig */
  int replicates;            /* permutation replicates */
  int include_observed;
  int has_seed;
  unsigned long long seed;
  int mode;                  /* kpc_exec_mode */
  int plain_test_limit;      /* KPC_PLAIN only: per-level cap on plain tests (0 = all) */
  /* KPC_PERM_R_STREAM replaces fastkpc's strict-route tables
   * (fastkpc/src/full_cuda_ci_one_call.cpp:464-567, drawn from R::unif_rand()
   * after the R facade's set.seed(seed), fastkpc/R/cuda_native.R:2291-2293):
   * dcc.perm draws energy dcov.test's index vector (one per test, mutated by
   * every replicate), hsic.perm draws sample(n) per replicate (R >= 3.6
   * "Rejection" sampling), and the variable permuted is the test's y. */
  int perm_stream;           /* kpc_perm_stream */
  /* KPC_PERM_R_STREAM: the stream to start from as .Random.seed[2:626]
   * (position, then the 624 state words); NULL = set.seed(seed) */
  const int* r_state;
  /* 1: consult and fill the process-wide result and target-state caches
   * (the reference's one-call route always does: CompactResultCache and
   * TargetStateCache, fastkpc/src/full_cuda_ci_one_call.cpp:1443-1755); a
   * replayed skeleton then runs no test.  0 (default): neither.  R-stream
   * permutation results are never cached (their indices depend on the
   * stream), their fits are. */
  int use_cache;
} kpc_options;

typedef struct kpc_result {
  int p;
  int levels;
  unsigned char* adjacency; /* p x p, row-major, 1 = edge kept */
  double* pmax;             /* p x p */
  int* sepset_len;          /* p x p; -1 = no sepset */
  int* sepset;              /* p x p x 64, 0-based variable indices */
  int n_edgetests[64];
  int n_tasks;              /* canonical trace rows (consumed tests) */
  int* task_level;
  int* task_x;
  int* task_y;
  int* task_slen;
  int* task_S;              /* n_tasks x 64 */
  double* task_p;
  unsigned char* task_deleted;
  double wall_seconds;
  double stage_seconds[8];  /* smooth, geometry, gcv, component, pair, host, h2d/d2h,
                               [7] whole-wave graphs (KPC_GRAPH=1: stages 0-4 not split) */
  long long physical_tests;
  long long residual_fits;
  long long components;
  long long setups;
  long long magic_evaluations;
  int truncated;            /* unused (kept for layout stability) */
  /* per-level counters (level = |S|) */
  long long level_tests[64];        /* consumed tests (== n_edgetests) */
  long long level_physical[64];     /* evaluated tests incl. speculation */
  long long level_fits[64];         /* GCV fits (unique (target, S)) */
  long long level_components[64];   /* dcov components */
  long long level_evaluations[64];  /* GCV objective evaluations */
  long long level_lanczos_steps[64];/* Lanczos steps summed over components */
  double level_seconds[64];         /* wall time of the level */
  /* KPC_PLAIN sampling: first plain_test_limit tests of each level run one
   * launch sequence per test (timed here); the rest run batched */
  long long plain_tests[64];
  double plain_seconds[64];
  long long stage_launches[8];      /* kernel launches per stage (as stage_seconds) */
  /* per trace row: smallest normalised decision margin of the test -- GCV
   * optimiser decisions of its two fits, incomplete-Cholesky pivot/stop
   * decisions (hsic.*), |T_perm - T| / |T| over the replicates (*.perm);
   * 1e300 when none applies.  Below ~1e-12 a rounding difference can flip a
   * discrete decision. */
  double* task_margin;
  /* device seconds per level and stage (as stage_seconds[0..4]) */
  double level_stage_seconds[64][5];
  /* per trace row, the conditioning set as 1-based indices joined by '|'
   * (the reference trace's S_key; "" for the empty set), rows separated by
   * '\n', NUL-terminated */
  char* task_keys;
  /* per trace row, the test statistic and the mean and variance of its null
   * distribution (gamma approximation; over the replicates for *.perm); the
   * reference trace carries them as method_statistic / method_mean /
   * method_variance for hsic.gamma
   * (fastkpc/src/full_cuda_ci_one_call.cpp:5620-5627) */
  double* task_stat;
  double* task_mean;
  double* task_var;
  /* per level, the tasks of the level plan (every conditioning set of both
   * orientations of every edge present at the level's start); the
   * reference's levels data frame column tasks_planned
   * (fastkpc/src/full_cuda_ci_one_call.cpp:5573, :5634) */
  long long level_tasks_planned[64];
  /* per trace row, the 1-based position of the test in its level's plan
   * (edges x < y in row-major order, the x orientation's conditioning sets
   * before the y orientation's, each in lexicographic order); the reference
   * trace's task_index (fastkpc/src/full_cuda_ci_one_call.cpp:5537) */
  int* task_index;
  /* per level, the planned tasks left on each edge after the test that
   * deleted it, summed (the levels data frame column
   * tasks_ignored_after_delete, fastkpc/src/full_cuda_ci_one_call.cpp:5486-5488),
   * and the device waves the level took (its column rounds, :5637: the
   * reference's frontier batches) */
  long long level_tasks_ignored[64];
  long long level_rounds[64];
  /* per trace row, the seed the row's permutation indices were drawn from
   * (dcc.perm / hsic.perm; 0 for the other methods); the trace column
   * permutation_seed (fastkpc/src/full_cuda_ci_one_call.cpp:5613-5616) */
  unsigned long long* task_perm_seed;
  /* KPC_PERM_R_STREAM: the stream after the last test (.Random.seed[2:626]
   * layout), for the caller to store back as R's RNGScope does */
  int r_state_out[625];
  /* use_cache: tests answered by the result cache, fits replayed from the
   * target-state cache */
  long long result_cache_hits;
  long long target_cache_hits;
} kpc_result;

void kpc_default_options(kpc_options* opt);
int kpc_skeleton_run(const double* data, int n, int p, const kpc_options* opt, kpc_result** out);
void kpc_result_free(kpc_result* r);

/* One batch of CI tests (x_i, y_i | S_i), S given as count x s_len 0-based
 * indices (all tests share s_len).  out: count x 4 {p.value, statistic, mean,
 * variance}.  With KPC_PERM_R_STREAM the tests draw from the stream in batch
 * order, as consecutive kernelCItest calls in R would. */
int kpc_ci_batch(const double* data, int n, int p, const kpc_options* opt, int count,
                 const int* x, const int* y, int s_len, const int* S, double* out);
/* kpc_ci_batch that also returns the R stream after the batch (r_state_out:
 * 625 ints, .Random.seed[2:626]; untouched unless KPC_PERM_R_STREAM). */
int kpc_ci_batch_rs(const double* data, int n, int p, const kpc_options* opt, int count,
                    const int* x, const int* y, int s_len, const int* S, double* out,
                    int* r_state_out);

/* regrXonS residuals of `targets` on S (count x s_len).  resid: count x n;
 * info: count x 32 {status, score, rss, edf, iterations, evaluations,
 * step_failed, numerical_rank, log_sp[0..21], ..., [30] = smallest normalised
 * optimiser decision margin} */
int kpc_regrxons_batch(const double* data, int n, int p, int count, const int* targets,
                       int s_len, const int* S, double* resid, double* info);

/* regrXonS fits with per-fit conditioning sets of mixed size: fit i regresses
 * column targets[i] on the s_len[i] (1..62) distinct 0-based columns (in the
 * given order: the order of the GAM's smooth terms)
 * S[s_off[i] .. s_off[i] + s_len[i]).  Fits are grouped by |S| (design width
 * q = 10, 30 or 1 + 9|S|): one design batch and one GCV launch per group, the
 * data uploaded once.  resid: count x n; info: count x 32 as in
 * kpc_regrxons_batch.  Any p >= 2.  Replaces one call per fit of
 * kpcalg regrXonS (kpcalg/R/regrXonS.R:33-61) / the reference's per-cohort GCV
 * optimiser calls (mgcv_multi_penalty_gcv_capacity.hpp). */
int kpc_regrxons_mixed(const double* data, int n, int p, int count, const int* targets,
                       const int* s_len, const int* s_off, const int* S, double* resid,
                       double* info);

/* Device time of the last kpc_regrxons_batch / kpc_regrxons_mixed call on
 * this thread (HIP events on the library's stream): out[0] design setup
 * seconds (smooth setups + geometry), out[1] GCV kernel seconds, out[2] GCV
 * launches.  No counterpart in the reference (its timers are R-side). */
int kpc_last_fit_timing(double* out);

/* Launch-shape autotuner state, one line per tuned knob ("key value=<v>
 * <candidate>:<seconds per replicate> ..."), NUL-terminated into buf (cap
 * bytes); returns the full length.  The tuned knobs change only how work is
 * split over workgroups, never results.  No counterpart in the reference. */
int kpc_tuning_report(char* buf, int cap);

/* GCV score, gradient (m) and Hessian (m x m) in log(sp) at log_sp (count x m)
 * for target on S. */
int kpc_gcv_grad_hess(const double* data, int n, int p, int count, const int* targets, int s_len,
                      const int* S, const double* log_sp, double* score, double* grad,
                      double* hess);

/* dcov.gamma low-rank components of `count` vectors (n each): values count x k,
 * centred vectors count x n x k, dsum, moment. */
int kpc_dcov_components(const double* x, int n, int count, int k, double* values,
                        double* vectors, double* dsum, double* moment);

/* Allocation-free kpc_dcov_components.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (HIP graph capture,
 * serving loops that forbid hipMalloc).  x (n x count, column-major), values
 * (count x k), vectors (count x k x n: vector e of component i at
 * (i * k + e) * n), dsum and moment (count) are DEVICE pointers of the caller;
 * the work is enqueued on `stream` (a hipStream_t; NULL = the default stream)
 * and the call returns without synchronising, so the outputs are ready when
 * the stream reaches that point.  Working memory is a 64 MiB device array
 * per GPU held by the library until the process ends; count <= 65536.
 * Calls that overlap on one GPU share that array and race.  Results equal
 * kpc_dcov_components bit for bit. */
int kpc_dcov_components_noalloc(const double* x, int n, int count, int k, double* values,
                                double* vectors, double* dsum, double* moment, void* stream);

/* Seed of the permutation test ({x, y}, S) for ci_method: SplitMix64 finaliser
 * folded over (user_seed, ci_method, min(x,y), max(x,y), s_len, S ascending).
 * Host only. */
int kpc_task_seed(unsigned long long user_seed, int ci_method, int x, int y, const int* S,
                  int s_len, unsigned long long* seed);

/* The permutation indices the dcc.perm / hsic.perm kernels consume, generated
 * on the device: out is count x replicates x n (0-based); replicate r of seed s
 * is the Durstenfeld shuffle of 0..n-1 (i = n-1 .. 1, j = floor(u (i+1) / 2^64))
 * with u the 64-bit word (i & 1 ? w3:w2 : w1:w0) of Philox4x32-10 at key s and
 * counter {i >> 1, r, 0x4b504331, 0}. */
int kpc_permutation_table(int n, int replicates, int count, const unsigned long long* seeds,
                          int* out);

/* ---- kpcalg-authority WAN-PDAG orientation ----
 * Replaces fastkpc_orient_wanpdag_kpcalg_authority (fastkpc/R/fast_kpc.R:
 * 263-445), i.e. kpcalg's udag2wanpdag (kpcalg/R/udag2wanpdag.R:104-436) with
 * regrVonPS (kpcalg/R/regrvonps.R:17-39): the collider step, the generalised
 * transitive step and rules 1-3, each regrVonPS call one GPU fit of V on
 * c(S, parents(V)) (in that term order) and one batch of GPU CI tests of the
 * residual against the raw columns of S.  Permutation methods draw either the
 * counter-based indices (per call from (seed, method, V, sorted c(S,
 * parents)), test i keyed (0, 1 + i)) or R's stream (KPC_PERM_R_STREAM,
 * continuing from r_state, e.g. the skeleton's r_state_out, as kpcalg does
 * after pcalg's skeleton). */
typedef struct kpc_orient_options {
  int ci_method;            /* kpc_ci_method */
  double alpha;             /* regrVonPS rejects a residual test at p < alpha */
  int num_col;
  double hsic_sig;
  double index;
  int replicates;
  int include_observed;
  int has_seed;
  unsigned long long seed;
  int perm_stream;          /* kpc_perm_stream */
  const int* r_state;       /* 625 ints (.Random.seed[2:626]) or NULL = set.seed(seed) */
  int orient_collider;      /* udag2wanpdag orientCollider */
  int solve_confl;          /* udag2wanpdag solve.confl (2 marks a bidirected edge) */
  int rules[3];             /* rules 1-3 on/off */
} kpc_orient_options;

typedef struct kpc_orient_result {
  int p;
  int* pdag;                /* p x p row-major: pdag[i * p + j] = 1 for i -> j or i - j */
  /* regrVonPS calls in order: target V, the subset S, the parents of V at
   * the call, the number of residual tests with p < alpha */
  int n_calls;
  int* call_v;
  int* call_slen;
  int* call_S;              /* n_calls x 64, 0-based */
  int* call_plen;
  int* call_parents;        /* n_calls x 64, 0-based */
  int* call_reject;
  /* the residual tests in order: call index, the S member tested, and
   * {p, statistic, mean, variance} */
  int n_tests;
  int* test_call;
  int* test_other;
  double* test_out;         /* n_tests x 4 */
  int r_state_out[625];     /* KPC_PERM_R_STREAM permutation methods: the stream afterwards */
  double wall_seconds;
} kpc_orient_result;

void kpc_default_orient_options(kpc_orient_options* opt);
/* adjacency: p x p symmetric 0/1 (row-major, e.g. kpc_result.adjacency);
 * sepset_len / sepset laid out as the kpc_result fields: p x p (-1 = none) and
 * p x p x 64. */
int kpc_wanpdag_run(const double* data, int n, int p, const unsigned char* adjacency,
                    const int* sepset_len, const int* sepset, const kpc_orient_options* opt,
                    kpc_orient_result** out);
void kpc_orient_result_free(kpc_orient_result* r);

/* Cache control, the counterpart of C_full_cuda_ci_one_call_cache_control
 * and _cache_state (fastkpc/src/full_cuda_ci_one_call.cpp:6812-6900). */
enum kpc_cache_action {
  KPC_CACHE_INFO = 0,             /* statistics only (capacity must be -1) */
  KPC_CACHE_RESET = 1,            /* empty both caches, zero their counters, new epoch */
  KPC_CACHE_CONFIGURE = 2,        /* result-cache capacity, 1 .. 1048576 */
  KPC_CACHE_CONFIGURE_TARGET = 3  /* target-state capacity, 1 .. 524288 */
};
typedef struct kpc_cache_info {
  long long capacity, entries, dataset_entries, requests, hits, misses, inserts, evictions,
      generation;
  long long target_capacity, target_entries, target_dataset_entries, target_requests,
      target_hits, target_misses, target_inserts, target_evictions, target_generation;
  long long epoch;
  char dataset_key[33]; /* kpc_cache_state: the data's 128-bit digest, hex */
} kpc_cache_info;
int kpc_cache_control(int action, long long capacity, kpc_cache_info* out);
/* Statistics plus the entries that belong to `data` (n x p, column-major). */
int kpc_cache_state(const double* data, int n, int p, kpc_cache_info* out);

/* Frees the device blocks the library keeps cached between calls (its arenas
 * are reused by later calls of the same thread instead of being freed).
 * Like torch.cuda.empty_cache; no counterpart in the reference. */
int kpc_empty_cache(void);

const char* kpc_last_error(void);
int kpc_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
