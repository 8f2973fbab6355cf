// Host accuracy check of stan_amd/include/stanx/sx_math.hpp against the C
// library: prints the largest error in ulps of each function over its
// domain (random and edge points).  Run by tests/test_host_cpp.py.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

#include "../../stan_amd/include/stanx/sx_math.hpp"

using namespace stanx::math;

static double ulps(double got, double ref) {
  if (got == ref) return 0.0;
  if (ref == 0.0) return std::fabs(got) / 4.9e-324;
  int e;
  std::frexp(ref, &e);
  const double ulp = std::fmax(std::ldexp(1.0, e - 53), 4.9406564584124654e-324);
  return std::fabs(got - ref) / ulp;
}

int main() {
  std::mt19937_64 g(12345);
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  double max_exp = 0, max_log = 0, max_rcp = 0, max_term = 0, max_r = 0;
  for (int i = 0; i < 2000000; ++i) {
    // exp on [-745, 0]: log-spread magnitudes
    const double x = -std::pow(10.0, -12.0 + 14.87 * u01(g));
    const double ref = std::exp(x);
    if (ref > 2.3e-308) max_exp = std::fmax(max_exp, ulps(sx_exp_nonpos(x), ref));
    const double t = (i & 1) ? u01(g) : std::pow(10.0, -300.0 * u01(g));
    max_log = std::fmax(max_log, ulps(sx_log1p_unit(t), std::log1p(t)));
    const double d = 1.0 + 2.0 * u01(g);
    max_rcp = std::fmax(max_rcp, ulps(sx_rcp(d), 1.0 / d));
    // composed row term vs a reference evaluation
    const double eta = (u01(g) - 0.5) * 80.0;
    const double s = (i & 2) ? 1.0 : -1.0;
    double term, r;
    sx_logit_row(eta, s, &term, &r);
    const double se = s * eta;
    const double rterm = -(se < 0 ? -se + std::log1p(std::exp(se)) : std::log1p(std::exp(-se)));
    const double rr = s / (1.0 + std::exp(se));
    max_term = std::fmax(max_term, std::fabs(term - rterm) / std::fmax(std::fabs(rterm), 1e-300));
    max_r = std::fmax(max_r, std::fabs(r - rr) / std::fmax(std::fabs(rr), 1e-300));
  }
  const double edge[] = {0.0, -0.0, -1e-320, -0.34657359027997264, -0.3465735902799727,
                         -708.0, -744.0, -745.1};
  for (double x : edge) {
    const double ref = std::exp(x);
    if (ref > 2.3e-308) max_exp = std::fmax(max_exp, ulps(sx_exp_nonpos(x), ref));
  }
  const double tedge[] = {0.0, 1.0, 0.41421356237309503, 0.414213562373095, 1e-300, 5e-324};
  for (double t : tedge) max_log = std::fmax(max_log, ulps(sx_log1p_unit(t), std::log1p(t)));
  std::printf("exp_ulps %.3f log1p_ulps %.3f rcp_ulps %.3f term_rel %.3e r_rel %.3e\n", max_exp,
              max_log, max_rcp, max_term, max_r);
  return 0;
}
