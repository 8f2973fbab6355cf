// Lean FP64 elementary functions for the logistic link, usable on host and
// device.  They cover only the ranges the link needs and trade the last
// half-ulp of the general library routines for a short instruction stream
// (the GPU pass evaluates them once per row of X, redundantly on every lane
// of a row).  Accuracy is checked on the host against the C library
// (a few ulp over each domain).
//
//   sx_exp_nonpos(x), x <= 0:
//       Cody and Waite (1980) reduction x = k ln2 + r, |r| <= ln2/2, with
//       ln2 split into a 32-bit head and a tail so that k * head is exact;
//       e^r from its Taylor series to degree 13 (truncation < 2e-17
//       relative); result scaled by 2^k.
//   sx_log1p_unit(t), 0 <= t <= 1:
//       log(1 + t) = 2 atanh(s), s = t / (2 + t), from the series
//       s + s^3/3 + s^5/5 + ... ; for t > sqrt(2) - 1 the argument is
//       halved, log(1 + t) = ln2 + 2 atanh((t - 1)/(t + 3)), so |s| <= 0.1716
//       and eleven terms reach double precision.
//   sx_rcp(d), 1 <= d <= 3: hardware reciprocal estimate refined by two
//       Newton-Raphson steps r <- r + r (1 - d r).
#ifndef STANX_SX_MATH_HPP
#define STANX_SX_MATH_HPP

#include <cmath>

#if defined(__HIPCC__)
#define SX_MATH_HD __host__ __device__ __forceinline__
#else
#define SX_MATH_HD inline
#endif

namespace stanx {
namespace math {

constexpr double kLn2 = 0x1.62e42fefa39efp-1;
constexpr double kLn2Hi = 0x1.62e42fee00000p-1;  // 32 significant bits
constexpr double kLn2Lo = 0x1.a39ef35793c76p-33; // ln2 - kLn2Hi
constexpr double kInvLn2 = 0x1.71547652b82fep+0;

SX_MATH_HD double sx_exp_nonpos(double x) {
  if (x < -745.2) return 0.0;
  const double k = rint(x * kInvLn2);
  double r = fma(-k, kLn2Hi, x);
  r = fma(-k, kLn2Lo, r);
  // Taylor series of e^r, Horner form, coefficients 1/n!
  double p = 1.0 / 6227020800.0;  // 1/13!
  p = fma(p, r, 1.0 / 479001600.0);
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, static_cast<int>(k));
}

SX_MATH_HD double sx_rcp(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(d);
#else
  double r = static_cast<double>(1.0f / static_cast<float>(d));
#endif
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

SX_MATH_HD double sx_log1p_unit(double t) {
  const bool high = t > 0x1.a827999fcef32p-2;  // sqrt(2) - 1
  const double num = high ? t - 1.0 : t;        // exact (Sterbenz) when high
  const double den = high ? t + 3.0 : t + 2.0;
  const double s = num * sx_rcp(den);
  const double z = s * s;
  // atanh(s)/s = sum_k z^k / (2k + 1)
  double p = 1.0 / 21.0;
  p = fma(p, z, 1.0 / 19.0);
  p = fma(p, z, 1.0 / 17.0);
  p = fma(p, z, 1.0 / 15.0);
  p = fma(p, z, 1.0 / 13.0);
  p = fma(p, z, 1.0 / 11.0);
  p = fma(p, z, 1.0 / 9.0);
  p = fma(p, z, 1.0 / 7.0);
  p = fma(p, z, 1.0 / 5.0);
  p = fma(p, z, 1.0 / 3.0);
  const double tail = (2.0 * s) * (z * p);  // 2 atanh(s) - 2 s
  if (!high) return fma(2.0, s, tail);
  return kLn2Hi + (fma(2.0, s, tail) + kLn2Lo);
}

// Bernoulli-logit row term and residual from the linear predictor eta and
// the sign s = 2y - 1:  term = -log(1 + e^{-s eta}), r = s inv_logit(-s eta)
// = y - inv_logit(eta), via log(1 + e^x) = max(x, 0) + log1p(e^{-|x|}).
SX_MATH_HD void sx_logit_row(double eta, double s, double* term, double* r) {
  const double se = s * eta;
  const double t = sx_exp_nonpos(-fabs(se));
  const double inv = sx_rcp(1.0 + t);
  *term = -(fmax(-se, 0.0) + sx_log1p_unit(t));
  *r = s * (se >= 0.0 ? t * inv : inv);
}

}  // namespace math
}  // namespace stanx
#endif
