// Launch interfaces of the gfx950 kernels (host side of the device code).
#pragma once
#include <hip/hip_runtime.h>

namespace kpc {

constexpr int kMagicInfo = 32;
// Per-test output record of every pair kernel: {p, statistic, mean, variance,
// decision margin} (margin 1e300 when the statistic has no discrete decision).
constexpr int kPairOut = 5;

struct DesignLayout {
  long long X, QR, tau, R, piv, rS, lsp0, norms, status, stride;
};

struct GeometryItem {
  int design_slot;
  int smooth_slot[64];
};

struct MagicItem {
  int design_slot;
  int target;
  int resid_slot;
  int pad;
};

struct MagicControl {
  double tol;
  int max_half_single;
  int max_half_multi;
  int max_iterations;
  double max_newton_step;
  double probe_step;
  int max_probes;
  double rank_tol;
};

DesignLayout host_design_layout(int n, int q, int m, int rk);

long long dcov_component_scratch_doubles(int n, int k);
size_t dcov_component_lds_bytes(int n);
hipError_t launch_dcov_components(const double* src_base, const long long* src_off,
                                  const int* dst_slot, int count, int n, int k, double* vals,
                                  double* vecs, double* dsum, double* moment, double* scratch,
                                  int grid, hipStream_t stream);
hipError_t launch_dcov_pairs(const int* px, const int* py, int count, int n, int k,
                             const double* vals, const double* vecs, const double* dsum,
                             const double* moment, double* out, int grid, hipStream_t stream);

long long smooth_scratch_doubles(int n);
size_t smooth_lds_bytes(int n);
hipError_t launch_smooth_setup(const double* data, int n, int p, const int* items, int count,
                               double* out_X, double* out_S, int* out_info, double* scratch,
                               int grid, hipStream_t stream);

long long geometry_scratch_doubles(int q);
// status_out[design_slot] repeats each design's status word in a compact
// array (one contiguous read for the host).
hipError_t launch_geometry(const GeometryItem* items, int count, int n, int q, int m, int dim,
                           int rk, const double* smooth_X, const double* smooth_S, double* design,
                           double* status_out, long long smooth_x_stride, double* scratch,
                           int grid, hipStream_t stream);

long long magic_scratch_doubles(int q, int m, int rk, int n);
size_t magic_lds_bytes(int n, int q, int m, bool aug_in_lds);
bool magic_aug_fits_lds(int n, int q, int m);
// fixed_info: fit records (kMagicInfo doubles each) of cached fits; an item
// with pad > 0 replays the final fit at the log(sp) of record pad - 1 and
// writes that record as its info (pad == 0: the full optimiser)
hipError_t launch_magic(const MagicItem* items, int count, int n, int q, int m, int rk,
                        const double* data, const double* design, double* resid, double* info,
                        double* scratch, int grid, const MagicControl& ctl, hipStream_t stream,
                        const double* fixed_info = nullptr);

}  // namespace kpc

namespace kpc {
hipError_t launch_gcv_eval(const MagicItem* items, int count, int n, int q, int m, int rk,
                           const double* data, const double* design, const double* lsp,
                           double* out, double* scratch, int grid, double rank_tol,
                           hipStream_t stream);
}  // namespace kpc

namespace kpc {
size_t dcov_lanczos_lds_bytes(int n, int k, int mmax);
long long dcov_lanczos_scratch_doubles(int n, int k, int mmax);
hipError_t launch_dcov_lanczos(const double* src_base, const long long* src_off,
                               const int* dst_slot, int count, int n, int k, double* vals,
                               double* vecs, double* dsum, double* moment, int* status,
                               double* scratch, int grid, int mmax, double tol,
                               hipStream_t stream);
int lanczos_mmax(int n);
}  // namespace kpc

namespace kpc {
// HSIC (kernlab inchol factors) and the permutation tests.
hipError_t launch_inchol_components(const double* src_base, const long long* src_off,
                                    const int* dst_slot, int count, int n, int k, double sigma,
                                    double* vecs, double* offdiag, double* moment, int* rank,
                                    double* margin, double* rows, int kp, int grid,
                                    hipStream_t stream);
hipError_t launch_hsic_pairs(const int* pa, const int* pb, int count, int n, int k,
                             const double* vecs, const int* rank, const double* offdiag,
                             const double* moment, double* out, int grid, hipStream_t stream);
hipError_t launch_dcc_rowmeans(const double* src_base, const long long* src_off,
                               const int* dst_slot, int count, int n, int stride, double index,
                               double* vecs, double* grand, int grid, hipStream_t stream);

// Permutation tests (kpc_perm.hip).  A unit is `nr` consecutive replicates
// r0 .. r0 + nr of test `test`; its permutations and values live at slots
// slot .. slot + nr of the chunk's scratch.
struct PermUnit {
  int test;
  int r0;
  int nr;
  int slot;
};
int perm_generate_width(int n);     // largest generation width for n
size_t dcc_perm_lds_bytes(int n, int rb, int jt);
bool dcc_perm_shape_supported(int jt, int rb);
// width: replicates per generation unit (<= perm_generate_width(n); each
// unit's lane shuffles one replicate in width * n * 2 bytes of LDS)
hipError_t launch_perm_generate(const PermUnit* units, int nunits,
                                const unsigned long long* seeds, int n, int width,
                                unsigned short* perms, int grid, hipStream_t stream);
// perms == nullptr: identity permutation (the observed statistic)
hipError_t launch_hsic_perm_values(const PermUnit* units, int nunits, const int* pa,
                                   const int* pb, int n, int kt, const double* rows,
                                   const unsigned short* perms, double* vals, int grid,
                                   hipStream_t stream);
hipError_t launch_dcc_perm_values(const PermUnit* units, int nunits, const int* pa, const int* pb,
                                  int n, int stride, double index, const double* vecs,
                                  const double* grand, const unsigned short* perms, double* vals,
                                  int jt, int rb, int grid, hipStream_t stream);
hipError_t launch_perm_tally(const PermUnit* segs, int nsegs, const double* vals,
                             const double* obs, double scale, double* state, hipStream_t stream);
hipError_t launch_perm_finalize(int count, const double* obs, const double* state, double scale,
                                int replicates, int include_observed, int one_if_nonpositive,
                                double* out, hipStream_t stream);
// Device-generated permutation table (the indices the permutation kernels
// consume), for verification against the oracle.
hipError_t launch_permutation_table(const unsigned long long* seeds, int count, int n,
                                    int replicates, int* out, hipStream_t stream);
}  // namespace kpc
