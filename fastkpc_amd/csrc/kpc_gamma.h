// Gamma-approximation tail shared by the dcov.gamma and hsic.gamma kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace kpc {

// ---------------------------------------------------------------- pgamma
// Lower regularised incomplete gamma P(a, x): power series for x < a + 1,
// otherwise 1 - Q with Q from the Legendre continued fraction (modified
// Lentz).  Same method as the oracle; kpcalg then reports 1 - P
// (kpcalg/R/dcovgamma.R:82).
__device__ inline double pgamma_lower(double x, double a) {
  if (!(a > 0.0) || isnan(x)) return __builtin_nan("");
  if (x <= 0.0) return 0.0;
  if (isinf(x)) return 1.0;
  const double lpre = -x + a * log(x) - lgamma(a);
  if (x < a + 1.0) {
    double ap = a, del = 1.0 / a, sum = del;
    for (int i = 0; i < 100000; ++i) {
      ap += 1.0;
      del *= x / ap;
      sum += del;
      if (fabs(del) < fabs(sum) * 1e-17) break;
    }
    return sum * exp(lpre);
  }
  const double tiny = 1e-300;
  double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int i = 1; i < 100000; ++i) {
    double an = -i * (i - a);
    b += 2.0;
    d = an * d + b;
    if (fabs(d) < tiny) d = tiny;
    c = b + an / c;
    if (fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < 1e-17) break;
  }
  return 1.0 - exp(lpre) * h;
}

__device__ inline double gamma_pvalue(double stat, double mean, double var) {
  const double alpha = (mean * mean) / var;
  const double beta = var / mean;
  const double rate = 1.0 / beta;
  const double scale = 1.0 / rate;
  return 1.0 - pgamma_lower(stat / scale, alpha);
}

}  // namespace kpc
