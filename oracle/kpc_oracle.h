/*
 * CPU oracle for the kernel-PC conditional-independence path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain serial C, the
 * reference algorithm (kpcalg 1.0.1 + mgcv 1.9-1 + pcalg stable skeleton, as
 * replicated by fastkpc's C++ sources) so that tests, the smoke check and the
 * benchmark's cpu_baseline leg can check the HIP product against it.  The
 * product library never links, loads or calls it.
 */
#ifndef KPC_ORACLE_H
#define KPC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- dcov.gamma (kpcalg/R/dcovgamma.R:56-82) ---- */
double orc_pgamma_lower(double x, double a);
double orc_gamma_pvalue(double stat, double mean, double var);
int orc_dcov_component(const double* x, int n, int k, double* values, double* centered,
                       double* dsum, double* moment);
double orc_weighted_cross_sum(const double* cx, const double* lx, const double* cy,
                              const double* ly, int n, int k);
void orc_dcov_pair_from_components(int n, int k, const double* cx, const double* lx,
                                   double dsum_x, double mom_x, const double* cy,
                                   const double* ly, double dsum_y, double mom_y,
                                   double out[5]);
/* out = {p.value, nV2, mean, variance, statistic} */
int orc_dcov_gamma(const double* x, const double* y, int n, int num_col, double out[5]);

/* ---- HSIC (kpcalg/R/hsicgamma.R:57-83, hsicperm.R:61-88; kernlab inchol) ---- */
int orc_inchol(const double* x, int n, double sigma, double tol, int maxr, double* g,
               int* pivots, double* margin);
int orc_hsic_component(const double* x, int n, double sig, int maxr, double* gc, int* rank,
                       double* offdiag, double* moment, double* margin);
double orc_hsic_cross(const double* gx, int rx, const double* gy, int ry, int n,
                      const int* perm);
/* out = {p.value, HSIC, mean, variance, margin} */
void orc_hsic_gamma_from_factors(const double* gx, int rx, double offx, double momx,
                                 const double* gy, int ry, double offy, double momy, int n,
                                 double out[5]);
void orc_hsic_perm_from_factors(const double* gx, int rx, const double* gy, int ry, int n,
                                int R, int include_observed, unsigned long long seed,
                                double out[5]);

/* ---- permutations (counter-based; shared with the GPU) and dcc.perm ---- */
void orc_philox4x32_10(const unsigned int ctr[4], const unsigned int key[2], unsigned int out[4]);
unsigned long long orc_task_seed(unsigned long long user_seed, int method, int x, int y,
                                 const int* S, int s_len);
void orc_permutation(unsigned long long seed, int replicate, int n, int* perm);
/* out = {p.value, n V^2, mean(rep), var(rep), margin} */
int orc_dcc_perm(const double* x, const double* y, int n, double index, int R,
                 int include_observed, unsigned long long seed, double out[5]);
/* the same with an explicit R x n index table (y permuted as y[table[r][i]]) */
int orc_dcc_perm_table(const double* x, const double* y, int n, double index, int R,
                       int include_observed, unsigned long long seed, const int* table,
                       double out[5]);
void orc_hsic_perm_table(const double* gx, int rx, const double* gy, int ry, int n, int R,
                         int include_observed, unsigned long long seed, const int* table,
                         double out[5]);

/* ---- R's global stream (Mersenne-Twister, set.seed) and kpcalg's
 *      permutation draws (fastkpc/src/full_cuda_ci_one_call.cpp:464-567) ---- */
typedef struct {
  unsigned int mt[624];
  int mti;
} orc_rng;
void orc_rng_set_seed(orc_rng* r, unsigned int seed);
unsigned int orc_rng_word(orc_rng* r);
double orc_unif_rand(orc_rng* r);
/* R_unif_index(dn): rejection ("Rejection" sample.kind) or floor(dn u) */
double orc_unif_index(orc_rng* r, double dn, int rounding);
/* sample(n), 0-based; pool: n ints of scratch */
void orc_r_sample(orc_rng* r, int n, int rounding, int* pool, int* out);
/* energy's permute(J, n) */
void orc_energy_permute(orc_rng* r, int* J, int n);
/* the R x n index table one permutation test draws (method ORC_DCC_PERM:
 * dcov.test; ORC_HSIC_PERM: hsic.perm); pool: n ints of scratch */
void orc_r_perm_table(orc_rng* r, int method, int n, int R, int* pool, int* table);

/* ---- regrXonS (kpcalg/R/regrXonS.R:33-61, mgcv gam GCV.Cp) ---- */
typedef struct {
  int n, q, m;  /* rows, coefficients, penalties */
  double* X;    /* n x q column-major, first column the intercept */
  double** S;   /* m penalty blocks, dim[i] x dim[i] column-major */
  int* off;     /* 0-based first coefficient of each block */
  int* dim;
  int* rank;
} orc_setup;

typedef struct {
  double tol;
  int max_step_halving_single;
  int max_step_halving_multi;
  int max_iterations;
  double max_newton_step;
  double boundary_probe_step;
  int max_boundary_probes;
  double rank_tol;
} orc_magic_control;

typedef struct {
  double score, rss, edf;
  double log_sp[64];
  int iterations, evaluations, step_failed;
  double margin; /* smallest normalised decision margin of the optimizer */
} orc_magic_result;

void orc_magic_default_control(orc_magic_control* c);
int orc_regr_setup(const double* data, int n, int ld, const int* S, int s_len, orc_setup* out);
void orc_setup_free(orc_setup* s);
int orc_magic_fit(const orc_setup* s, const double* y, const orc_magic_control* ctl,
                  double* resid, orc_magic_result* res);
int orc_magic_eval(const orc_setup* s, const double* y, const double* lsp, double rank_tol,
                   double* score, double* grad, double* hess);
int orc_regrxons(const double* data, int n, int ld, int target, const int* S, int s_len,
                 const orc_magic_control* ctl, double* resid, orc_magic_result* res);
/* The final fit of orc_regrxons at a given log(sp) instead of the optimiser's
 * choice: residuals and the GCV score / rss / edf there.  Used to replay a
 * fit whose optimiser took a rounding-decided step at the end point another
 * implementation selected. */
int orc_regrxons_at_sp(const double* data, int n, int ld, int target, const int* S, int s_len,
                       const double* log_sp, double rank_tol, double* resid, double out[3]);

/* ---- pcalg stable skeleton (fastkpc/src/skeleton_engine.cpp:96-227) ---- */
enum { ORC_DCC_GAMMA = 0, ORC_HSIC_GAMMA = 1, ORC_DCC_PERM = 2, ORC_HSIC_PERM = 3 };

typedef struct {
  int ci_method;
  double alpha;
  int m_max;        /* resolved maximum conditioning size */
  int num_col;
  double hsic_sig;
  int replicates;   /* permutation methods */
  unsigned long long perm_seed;
  int threads;      /* OpenMP threads for the per-level precompute (0 = serial) */
  double index;     /* dcc.perm distance exponent */
  int include_observed;
  /* 1: permutation indices from R's global stream after set.seed(perm_seed),
   * tests run in pcalg's order (orientations by x, then y); 0: the
   * counter-based indices of the task seed */
  int r_stream;
} orc_skeleton_options;

typedef struct {
  int p, levels;
  unsigned char* adjacency; /* p x p row-major */
  double* pmax;             /* p x p row-major */
  int* sepset_len;          /* p x p; -1 = none */
  int* sepset;              /* p x p x 64 */
  int n_edgetests[64];
  int trace_len, trace_cap;
  int* trace_level; int* trace_x; int* trace_y; int* trace_slen; int* trace_S; /* S: cap x 64 */
  double* trace_p;
  unsigned char* trace_del;
  double* trace_margin; /* min decision margin of the test's two fits */
  /* per trace row: statistic and null mean / variance (dcc.gamma: nV^2 and
   * its gamma moments; hsic.gamma: HSIC and its moments; *.perm: the observed
   * statistic and the replicates' mean / variance) */
  double* trace_stat; double* trace_mean; double* trace_var;
  orc_rng rng_out;      /* r_stream: the stream after the last test */
} orc_skeleton_result;

int orc_skeleton(const double* data, int n, int p, const orc_skeleton_options* opt,
                 orc_skeleton_result* res);
void orc_skeleton_result_free(orc_skeleton_result* r);

/* ---- kpcalg WAN-PDAG orientation (kpcalg/R/udag2wanpdag.R,
 *      kpcalg/R/regrvonps.R; fastkpc/R/fast_kpc.R:263-445) ---- */
typedef struct {
  int ci_method;
  double alpha;
  int num_col;
  double hsic_sig;
  int replicates;
  int include_observed;
  double index;
  unsigned long long perm_seed; /* counter-based indices: orc_orient_seed(perm_seed, ...) */
  int r_stream;                 /* 1: R's stream, from r_state_in or set.seed(perm_seed) */
  const orc_rng* r_state_in;
  int orient_collider, solve_confl;
  int rules[3];
} orc_orient_options;

typedef struct {
  int* pdag; /* p x p row-major: 1 / 0 (2 = bidirected with solve_confl) */
  int n_calls, call_cap;
  int* call_v; int* call_slen; int* call_S; int* call_plen; int* call_parents; int* call_reject;
  int n_tests, test_cap;
  int* test_call; int* test_other;
  double* test_out; /* n_tests x 4 {p, statistic, mean, variance} */
  orc_rng rng_out;
} orc_orient_result;

unsigned long long orc_orient_seed(unsigned long long user_seed, int method, int V, int i,
                                   const int* cond, int c_len);
int orc_wanpdag(const double* data, int n, int p, const unsigned char* adjacency,
                const int* sepset_len, const int* sepset, const orc_orient_options* opt,
                orc_orient_result* res);
void orc_orient_result_free(orc_orient_result* r);

/* single CI test (kernelCItest semantics):
 * out = {p.value, statistic, mean, variance, decision margin} */
int orc_ci_test(const double* data, int n, int p, int x, int y, const int* S, int s_len,
                const orc_skeleton_options* opt, double out[5]);

#ifdef __cplusplus
}
#endif
#endif
