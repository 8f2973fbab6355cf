// Device helpers shared by the GLM passes: wave reductions and the final
// assembly of log density and gradient from the reduced statistics
// st[0] = sum of row terms, st[1] = sum of residuals, st[2..K+1] = X^T r,
// st[K+2] = count of non-finite linear predictors.
#ifndef STANX_GLM_FINALIZE_HPP
#define STANX_GLM_FINALIZE_HPP

#include <hip/hip_runtime.h>

#include "sx_internal.hpp"

namespace sx {
namespace {

constexpr double kHalfLog2Pi = 0.91893853320467274178032973640562;

// Sum of v over the 64 lanes (xor butterfly): every lane gets the total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Adds priors, Jacobian and scaling to the reduced statistics st[0..E) and
// writes out[0] = lp, out[1..D] = gradient, out[D+1] = status (0 ok,
// 1 domain error).  Runs on the first wave of the reduction launch.
template <int FAM>
__device__ void finalize(const pass_args& a, const double* st) {
  const int lane = threadIdx.x & 63;
  const int K = a.K;
  const double* th = a.theta;
  const double alpha = th[0];
  double lp_prior = 0.0;  // lane-partial of prior terms over beta
  double scale = 1.0;
  if (FAM == SX_NORMAL) {
    const double sigma = exp(th[K + 1]);
    scale = 1.0 / (sigma * sigma);
  }
  for (int k = lane; k < K; k += 64) {
    const double b = th[1 + k];
    double g = st[2 + k] * scale;
    if (a.beta_sd > 0) {
      const double z = b / a.beta_sd;
      lp_prior += -0.5 * z * z;
      if (!a.propto) lp_prior += a.beta_const;
      g += -b / (a.beta_sd * a.beta_sd);
    }
    a.out[2 + k] = g;
  }
  lp_prior = wave_sum(lp_prior);
  if (lane != 0) return;
  double lp = 0.0;
  double ga = st[1] * scale;
  double bad = st[K + 2];
  if (FAM == SX_LOGIT) {
    lp = st[0];
  } else {
    const double u = th[K + 1];
    const double sigma = exp(u);
    if (!(sigma > 0) || !isfinite(sigma)) bad += 1.0;
    const double SS = st[0];
    lp = -0.5 * SS * scale - static_cast<double>(a.N) * u;
    if (!a.propto) lp += -static_cast<double>(a.N) * kHalfLog2Pi;
    double gu = -static_cast<double>(a.N) + SS * scale;
    if (a.sigma_sd > 0) {
      // sigma ~ normal(sigma_mu, sigma_sd): d/du = -(sigma - mu) sigma / sd^2
      const double z = (sigma - a.sigma_mu) / a.sigma_sd;
      lp += -0.5 * z * z;
      if (!a.propto) lp += a.sigma_const;
      gu += -z * (sigma / a.sigma_sd);
    }
    if (a.jacobian) {
      lp += u;
      gu += 1.0;
    }
    a.out[K + 2] = gu;
  }
  if (a.alpha_sd > 0) {
    const double z = alpha / a.alpha_sd;
    lp += -0.5 * z * z;
    if (!a.propto) lp += a.alpha_const;
    ga += -alpha / (a.alpha_sd * a.alpha_sd);
  }
  lp += lp_prior;
  a.out[0] = lp;
  a.out[1] = ga;
  if (!isfinite(lp)) bad += 1.0;
  a.out[a.D + 1] = bad > 0 ? 1.0 : 0.0;
}

}  // namespace
}  // namespace sx
#endif
