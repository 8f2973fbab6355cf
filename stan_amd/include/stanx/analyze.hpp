// Posterior diagnostics on the host, restating (paths relative to the
// stan-dev/stan 2.35.0 tree, src/stan/analyze/mcmc/):
//   autocovariance             autocovariance.hpp:39-100 (FFT, biased 1/N)
//   ess                        ess.hpp:28-103 (Geyer initial monotone sequence)
//   split_chains               split_chains.hpp:49-67
//   rank_transform             rank_normalization.hpp:23-57 (average ranks,
//                              Blom offsets 3/8)
//   split_rank_normalized_ess  split_rank_normalized_ess.hpp:28-55 (bulk)
//   is_finite_and_varies       check_chains.hpp:20-38
// Chains are given column-major: element (iteration i, chain c) at
// x[c * n + i].  The FFT is an iterative radix-2 Cooley-Tukey transform on
// the zero-padded series (length a power of two >= 2n).
#ifndef STANX_ANALYZE_HPP
#define STANX_ANALYZE_HPP

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstddef>
#include <limits>
#include <numeric>
#include <vector>

namespace stanx {
namespace analyze {

namespace detail {

inline void fft(std::vector<std::complex<double>>& a, bool inverse) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  const double pi = 3.14159265358979323846;
  for (size_t len = 2; len <= n; len <<= 1) {
    const double ang = 2 * pi / static_cast<double>(len) * (inverse ? 1 : -1);
    const std::complex<double> wl(std::cos(ang), std::sin(ang));
    for (size_t i = 0; i < n; i += len) {
      std::complex<double> w(1.0, 0.0);
      for (size_t j = 0; j < len / 2; ++j) {
        const std::complex<double> u = a[i + j], v = a[i + j + len / 2] * w;
        a[i + j] = u + v;
        a[i + j + len / 2] = u - v;
        w *= wl;
      }
    }
  }
  if (inverse)
    for (auto& v : a) v /= static_cast<double>(n);
}

// Inverse of the standard normal CDF: P. J. Acklam's rational
// approximations (2003; relative error < 1.2e-9), refined by one Halley
// step on erfc, which brings it to a few ulp.
inline double inv_phi(double p) {
  if (p <= 0.0) return -std::numeric_limits<double>::infinity();
  if (p >= 1.0) return std::numeric_limits<double>::infinity();
  static const double a[6] = {-3.969683028665376e+01, 2.209460984245205e+02,
                              -2.759285104469687e+02, 1.383577518672690e+02,
                              -3.066479806614716e+01, 2.506628277459239e+00};
  static const double b[5] = {-5.447609879822406e+01, 1.615858368580409e+02,
                              -1.556989798598866e+02, 6.680131188771972e+01,
                              -1.328068155288572e+01};
  static const double c[6] = {-7.784894002430293e-03, -3.223964580411365e-01,
                              -2.400758277161838e+00, -2.549732539343734e+00,
                              4.374664141464968e+00, 2.938163982698783e+00};
  static const double d[4] = {7.784695709041462e-03, 3.224671290700398e-01,
                              2.445134137142996e+00, 3.754408661907416e+00};
  const double plow = 0.02425;
  double x;
  if (p < plow) {
    const double q = std::sqrt(-2 * std::log(p));
    x = (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
        ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1);
  } else if (p <= 1 - plow) {
    const double q = p - 0.5, r = q * q;
    x = (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
        (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1);
  } else {
    const double q = std::sqrt(-2 * std::log1p(-p));
    x = -(((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
        ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1);
  }
  // Halley step on f(x) = Phi(x) - p, written as (1 - p) - Q(x) in the
  // upper half, where 1 - p is exact
  const double e = p <= 0.5 ? 0.5 * std::erfc(-x / std::sqrt(2.0)) - p
                            : (1.0 - p) - 0.5 * std::erfc(x / std::sqrt(2.0));
  const double u = e * std::sqrt(2 * 3.14159265358979323846) * std::exp(x * x / 2);
  return x - u / (1 + x * u / 2);
}

// |v - ref| <= 1e-12 min(|v|, |ref|) for every element (Eigen isApproxToConstant)
inline bool approx_const(const double* v, size_t n, size_t stride, double ref) {
  for (size_t i = 0; i < n; ++i) {
    const double x = v[i * stride];
    if (!(std::fabs(x - ref) <= 1e-12 * std::min(std::fabs(x), std::fabs(ref)))) return false;
  }
  return true;
}

}  // namespace detail

// autocovariance of y[0..n): biased estimator, all lags
inline std::vector<double> autocovariance(const double* y, size_t n) {
  size_t M = 1;
  while (M < n) M <<= 1;
  double mean = 0.0;
  for (size_t i = 0; i < n; ++i) mean += y[i];
  mean /= static_cast<double>(n);
  std::vector<std::complex<double>> c(2 * M, 0.0);
  for (size_t i = 0; i < n; ++i) c[i] = y[i] - mean;
  detail::fft(c, false);
  for (auto& v : c) v = std::complex<double>(std::norm(v), 0.0);
  detail::fft(c, true);
  std::vector<double> ac(n);
  double var = 0.0;
  for (size_t i = 0; i < n; ++i) var += (y[i] - mean) * (y[i] - mean);
  var /= static_cast<double>(n);
  const double c0 = c[0].real();
  for (size_t i = 0; i < n; ++i) ac[i] = c[i].real() / c0 * var;
  return ac;
}

// is_finite_and_varies over chains (column-major n x m)
inline bool is_finite_and_varies(const double* x, size_t n, size_t m) {
  for (size_t i = 0; i < n * m; ++i)
    if (!std::isfinite(x[i])) return false;
  for (size_t c = 0; c < m; ++c)
    if (detail::approx_const(x + c * n, n, 1, x[c * n])) return false;
  if (m > 1 && detail::approx_const(x, m, n, x[0])) return false;
  return true;
}

// ess of column-major chains (n x m)
inline double ess(const double* x, size_t n, size_t m) {
  std::vector<std::vector<double>> acov(m);
  std::vector<double> chain_mean(m);
  for (size_t c = 0; c < m; ++c) {
    acov[c] = autocovariance(x + c * n, n);
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += x[c * n + i];
    chain_mean[c] = s / static_cast<double>(n);
  }
  double w = 0.0;
  for (size_t c = 0; c < m; ++c) w += acov[c][0] * n / (n - 1.0);
  w /= static_cast<double>(m);
  double var_plus = w * (n - 1.0) / n;
  if (m > 1) {
    double mu = 0.0;
    for (double v : chain_mean) mu += v;
    mu /= static_cast<double>(m);
    double ss = 0.0;
    for (double v : chain_mean) ss += (v - mu) * (v - mu);
    var_plus += ss / (m - 1.0);
  }
  auto acov_mean = [&](size_t t) {
    double s = 0.0;
    for (size_t c = 0; c < m; ++c) s += acov[c][t];
    return s / static_cast<double>(m);
  };
  std::vector<double> rho(n, 0.0);
  double rho_even = 1.0;
  rho[0] = rho_even;
  double rho_odd = 1 - (w - acov_mean(1)) / var_plus;
  rho[1] = rho_odd;
  size_t t = 1;
  while (t + 4 < n && (rho_even + rho_odd > 0) && !std::isnan(rho_even + rho_odd)) {
    rho_even = 1 - (w - acov_mean(t + 1)) / var_plus;
    rho_odd = 1 - (w - acov_mean(t + 2)) / var_plus;
    if (rho_even + rho_odd >= 0) {
      rho[t + 1] = rho_even;
      rho[t + 2] = rho_odd;
    }
    if (rho[t + 1] + rho[t + 2] > rho[t - 1] + rho[t]) {
      rho[t + 1] = (rho[t - 1] + rho[t]) / 2;
      rho[t + 2] = rho[t + 1];
    }
    t += 2;
  }
  const size_t max_t = t;
  if (rho_even > 0) rho[max_t + 1] = rho_even;
  const double total = static_cast<double>(m * n);
  double tau = -1;
  for (size_t i = 0; i < max_t; ++i) tau += 2 * rho[i];
  tau += rho[max_t + 1];
  tau = std::max(tau, 1 / std::log10(total));
  return total / tau;
}

// split each chain in halves (the middle draw of an odd chain is dropped)
inline std::vector<double> split_chains(const double* x, size_t n, size_t m) {
  const size_t half = n / 2, tail = (n + 1) / 2;
  std::vector<double> s(2 * m * half);
  for (size_t c = 0; c < m; ++c)
    for (size_t i = 0; i < half; ++i) {
      s[(2 * c) * half + i] = x[c * n + i];
      s[(2 * c + 1) * half + i] = x[c * n + tail + i];
    }
  return s;
}

// normal scores of the average ranks, same shape
inline std::vector<double> rank_transform(const double* x, size_t size) {
  std::vector<size_t> order(size);
  std::iota(order.begin(), order.end(), size_t(0));
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return x[a] < x[b]; });
  std::vector<double> z(size);
  size_t i = 0;
  while (i < size) {
    size_t j = i + 1;
    while (j < size && x[order[j]] == x[order[i]]) ++j;
    const double avg = (static_cast<double>(i + 1) + static_cast<double>(j)) / 2.0;
    // inverse normal CDF of (avg - 3/8) / (size + 1/4)
    const double p = (avg - 0.375) / (static_cast<double>(size) + 0.25);
    const double q = detail::inv_phi(p);
    for (size_t k = i; k < j; ++k) z[order[k]] = q;
    i = j;
  }
  return z;
}

// bulk ESS of split_rank_normalized_ess; NaN if the split chains do not
// vary or have fewer than 4 draws
inline double bulk_ess(const double* x, size_t n, size_t m) {
  const size_t half = n / 2;
  if (half < 4) return std::numeric_limits<double>::quiet_NaN();
  std::vector<double> s = split_chains(x, n, m);
  if (!is_finite_and_varies(s.data(), half, 2 * m)) return std::numeric_limits<double>::quiet_NaN();
  std::vector<double> z = rank_transform(s.data(), s.size());
  return ess(z.data(), half, 2 * m);
}

// minimum bulk ESS over parameters of draws laid out [chain][iteration][param
// stride], parameters p0 .. p0 + P - 1
inline double min_bulk_ess(const double* draws, size_t n, size_t m, size_t stride,
                           size_t chain_stride, size_t p0, size_t P) {
  double best = std::numeric_limits<double>::quiet_NaN();
  std::vector<double> x(n * m);
  for (size_t p = p0; p < p0 + P; ++p) {
    for (size_t c = 0; c < m; ++c)
      for (size_t i = 0; i < n; ++i) x[c * n + i] = draws[c * chain_stride + i * stride + p];
    const double e = bulk_ess(x.data(), n, m);
    if (std::isfinite(e) && !(e >= best)) best = e;
  }
  return best;
}

}  // namespace analyze
}  // namespace stanx
#endif
