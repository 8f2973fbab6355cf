// Host checks of the dense_e pieces of the C++ sampler, restating the
// reference's own tests:
//   covar_adaptation_test.cpp:5-28     zero-variance window -> 1e-3 5/15 I
//   dense_e_metric_test.cpp:11-59      sample_p covariance within 5 sigma
//                                      of the metric (Wishart spread)
//   dense_e_metric_test.cpp:62-150     dtau_dp / dphi_dq against central
//                                      finite differences (funnel, eps 1e-6)
//   derived_nuts_test.cpp:94-132       compute_criterion with a dense metric
// Prints "name value" lines; the pytest side states the tolerances.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../stan_amd/include/stanx/callbacks.hpp"
#include "../../stan_amd/include/stanx/mcmc/adapt_dense_e_nuts.hpp"
#include "../../stan_amd/include/stanx/model.hpp"

using stanx::mcmc::vec_t;

// funnel (src/test/test-models/good/mcmc/hmc/hamiltonians/funnel.stan)
static int funnel(void*, const double* q, int D, double* lp, double* g) {
  const double v = q[0];
  double l = -0.5 * v * v / 9.0, gv = -v / 9.0;
  const double e2 = std::exp(-2.0 * v);
  for (int i = 1; i < D; ++i) {
    l += -0.5 * q[i] * q[i] * e2 - v;
    gv += q[i] * q[i] * e2 - 1.0;
    g[i] = -q[i] * e2;
  }
  g[0] = gv;
  *lp = l;
  return 0;
}

int main() {
  stanx::callbacks::logger logger;
  // ---- covariance adaptation: 10 samples of q = 0, window of 10 ----
  {
    const int n = 10;
    stanx::mcmc::covar_adaptation adapter(n, "covariance");
    adapter.set_window_params(50, 0, 0, 10, logger);
    vec_t covar(n * n, 0.0), q(n, 0.0);
    for (int i = 0; i < 10; ++i) adapter.observe(covar, q);
    double worst = 0.0;
    const double target = 1e-3 * 5.0 / (10 + 5.0);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        worst = std::fmax(worst, std::fabs(covar[i * n + j] - (i == j ? target : 0.0)));
    std::printf("covar_adapt_maxdiff %.17g\n", worst);
  }
  stanx::model::callback_model model(funnel, nullptr, 2);
  // ---- sample_p: the sample covariance of p is the metric M (2x2) ----
  {
    const double m[4] = {3.0, -2.0, -2.0, 4.0};
    const double det = m[0] * m[3] - m[1] * m[2];
    vec_t minv = {m[3] / det, -m[1] / det, -m[2] / det, m[0] / det};
    stanx::mcmc::dense_e_metric<stanx::model::callback_model> metric(model);
    stanx::mcmc::dense_e_point z(2);
    z.set_metric(minv);
    stanx::rng_t rng(0, 0);
    const int ns = 1000;
    double c[4] = {0, 0, 0, 0};
    for (int s = 0; s < ns; ++s) {
      metric.sample_p(z, rng);
      c[0] += z.p[0] * z.p[0] / ns;
      c[1] += z.p[0] * z.p[1] / ns;
      c[2] += z.p[1] * z.p[0] / ns;
      c[3] += z.p[1] * z.p[1] / ns;
    }
    const double var[4] = {2 * m[0] * m[0], m[1] * m[1] + m[3] * m[0],
                           m[2] * m[2] + m[3] * m[0], 2 * m[3] * m[3]};
    double worst = 0.0;  // in units of the Wishart standard deviation
    for (int k = 0; k < 4; ++k) worst = std::fmax(worst, std::fabs(m[k] - c[k]) / std::sqrt(var[k] / ns));
    std::printf("sample_p_sigmas %.17g\n", worst);
  }
  // ---- finite differences of tau = T and phi = V (funnel, D = 11) ----
  {
    stanx::model::callback_model funnel11(funnel, nullptr, 11);
    stanx::mcmc::dense_e_metric<stanx::model::callback_model> metric(funnel11);
    stanx::mcmc::dense_e_point z(11);
    // a non-diagonal metric: identity plus 0.3 on the first off-diagonals
    vec_t minv(121, 0.0);
    for (int i = 0; i < 11; ++i) {
      minv[i * 11 + i] = 1.0;
      if (i + 1 < 11) minv[i * 11 + i + 1] = minv[(i + 1) * 11 + i] = 0.3;
    }
    z.set_metric(minv);
    for (int i = 0; i < 11; ++i) z.q[i] = z.p[i] = 1.0;
    metric.init(z, logger);
    const double eps = 1e-6;
    vec_t g2;
    metric.dtau_dp(z, g2);
    double e_tau = 0.0;
    for (int i = 0; i < 11; ++i) {
      z.p[i] += eps;
      double d = metric.T(z);
      z.p[i] -= 2 * eps;
      d -= metric.T(z);
      z.p[i] += eps;
      e_tau = std::fmax(e_tau, std::fabs(d / (2 * eps) - g2[i]));
    }
    const vec_t g3 = z.g;  // dphi_dq = g
    double e_phi = 0.0;
    for (int i = 0; i < 11; ++i) {
      z.q[i] += eps;
      metric.init(z, logger);
      double d = z.V;
      z.q[i] -= 2 * eps;
      metric.init(z, logger);
      d -= z.V;
      z.q[i] += eps;
      metric.init(z, logger);
      e_phi = std::fmax(e_phi, std::fabs(d / (2 * eps) - g3[i]));
    }
    std::printf("dtau_dp_fd %.17g\ndphi_dq_fd %.17g\n", e_tau, e_phi);
  }
  // ---- compute_criterion with a dense metric (model size 1) ----
  {
    using nuts = stanx::mcmc::dense_e_nuts<stanx::model::callback_model>;
    stanx::model::callback_model one(funnel, nullptr, 1);
    stanx::mcmc::dense_e_point start(1), finish(1);
    stanx::mcmc::dense_e_metric<stanx::model::callback_model> metric(one);
    vec_t ps0, ps1;
    start.q[0] = 1;
    start.p[0] = 1;
    finish.q[0] = 2;
    finish.p[0] = 1;
    metric.dtau_dp(start, ps0);
    metric.dtau_dp(finish, ps1);
    vec_t rho = {start.p[0] + finish.p[0]};
    const bool a = nuts::compute_criterion(ps0, ps1, rho);
    finish.p[0] = -1;
    metric.dtau_dp(finish, ps1);
    rho[0] = start.p[0] + finish.p[0];
    const bool b = nuts::compute_criterion(ps0, ps1, rho);
    std::printf("criterion_true %d\ncriterion_false %d\n", a ? 1 : 0, b ? 0 : 1);
  }
  return 0;
}
