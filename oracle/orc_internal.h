/* Internal declarations shared by the oracle's translation units. */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

int orc_pstf2_upper(double* a, int n, int lda, int* piv, double tol, int* rank);
void orc_larfg(int m, double* alpha, double* x, int incx, double* tau);
void orc_geqp3(double* a, int m, int n, int lda, int* jpvt, double* tau);
void orc_apply_qt(const double* a, int m, int k, int lda, const double* tau, double* y);
int orc_svd_jacobi(const double* a, int m, int n, double* u, double* s, double* v);
int orc_syev_jacobi(const double* a, int n, double* w, double* z);
int orc_tridiag_ql(double* d, double* e, int n, double* z, int ldz);
void orc_tridiag_reduce(double* a, int n, int lda, double* d, double* e, double* tau);
void orc_tridiag_backtransform(const double* a, int n, int lda, const double* tau,
                               double* zz, int k);
void orc_tridiag_invit(const double* d, const double* e, int n, double lam,
                       const double* prev, int nprev, double* z, unsigned* seed);

/* R-style mean: long double accumulation plus a correction pass. */
double orc_r_mean(const double* x, int n, int stride);

#endif
