// A consumer of the model boundary written the way a generated model is: a
// class deriving from model_base_crtp<M> with one log_prob template over
// double and the library's var (src/stan/model/model_base_crtp.hpp:73-301).
// Checks, printed as "name value" lines for the pytest side:
//   - the gradient the var tape gives against the closed form;
//   - the eight log_prob overloads (double and var, propto x jacobian)
//     against each other and the closed-form constants;
//   - a precomputed-gradient entry (how a GPU model enters a tape) inside
//     a larger expression;
//   - the tape is empty after a model throws;
//   - the adaptive service over the model: posterior mean and sd of mu.
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "../../stan_amd/include/stanx/callbacks.hpp"
#include "../../stan_amd/include/stanx/model.hpp"
#include "../../stan_amd/include/stanx/services.hpp"
#include "../../stan_amd/include/stanx/var.hpp"

using stanx::math::var;

// y ~ normal(mu, sigma); mu ~ normal(0, 10); sigma ~ normal(0, 5), sigma > 0
// (unconstrained u = log sigma, Jacobian + u).
class normal_model : public stanx::model::model_base_crtp<normal_model> {
 public:
  explicit normal_model(std::vector<double> y) : model_base_crtp(2), y_(std::move(y)) {}

  template <bool propto, bool jacobian, typename T>
  T log_prob(std::vector<T>& params_r, std::vector<int>& params_i,
             std::ostream* msgs = nullptr) const {
    using std::exp;
    using std::log;
    using stanx::math::exp;
    using stanx::math::square;
    const T mu = params_r[0];
    const T u = params_r[1];
    if (stanx::math::value_of(u) > 50) throw std::domain_error("sigma overflows");
    const T sigma = exp(u);
    T lp = 0.0;
    if (jacobian) lp += u;
    const double half_log_2pi = 0.5 * std::log(2 * M_PI);
    lp += -0.5 * square(mu / 10.0);
    lp += -0.5 * square(sigma / 5.0);
    if (!propto) lp += -std::log(10.0) - std::log(5.0) - 2 * half_log_2pi + std::log(2.0);
    for (double yn : y_) {
      lp += -0.5 * square((yn - mu) / sigma) - u;
      if (!propto) lp += -half_log_2pi;
    }
    return lp;
  }

  // closed form of the propto + jacobian density and its gradient
  double closed(const std::vector<double>& x, std::vector<double>& g) const {
    const double mu = x[0], u = x[1], s = std::exp(u);
    double lp = u - 0.5 * (mu / 10) * (mu / 10) - 0.5 * (s / 5) * (s / 5);
    double gm = -mu / 100.0, gu = 1.0 - s * s / 25.0;
    for (double yn : y_) {
      const double r = (yn - mu) / s;
      lp += -0.5 * r * r - u;
      gm += r / s;
      gu += r * r - 1.0;
    }
    g = {gm, gu};
    return lp;
  }

  std::string model_name() const override { return "normal_model"; }
  void get_param_names(std::vector<std::string>& n, bool, bool) const override {
    n = {"mu", "sigma"};
  }
  void get_dims(std::vector<std::vector<size_t>>& d, bool, bool) const override { d = {{}, {}}; }
  void constrained_param_names(std::vector<std::string>& n, bool, bool) const override {
    n.push_back("mu");
    n.push_back("sigma");
  }
  void unconstrained_param_names(std::vector<std::string>& n, bool, bool) const override {
    n.push_back("mu");
    n.push_back("sigma");
  }
  void transform_inits(const stanx::io::var_context& c, std::vector<int>&,
                       std::vector<double>& r, std::ostream*) const override {
    r = {c.vals_r("mu")[0], std::log(c.vals_r("sigma")[0])};
  }
  void write_array(stanx::rng_t&, std::vector<double>& r, std::vector<int>&,
                   std::vector<double>& vars, bool, bool, std::ostream*) const override {
    vars = {r[0], std::exp(r[1])};
  }
  void unconstrain_array(const std::vector<double>& c, std::vector<double>& u,
                         std::ostream*) const override {
    u = {c[0], std::log(c[1])};
  }

 private:
  std::vector<double> y_;
};

static int gauss(void*, const double* q, int D, double* lp, double* g) {
  double s = 0;
  for (int i = 0; i < D; ++i) {
    s += q[i] * q[i];
    g[i] = -q[i];
  }
  *lp = -0.5 * s;
  return 0;
}

int main() {
  std::vector<double> y;
  for (int n = 0; n < 100; ++n) y.push_back(3.0 + 2.0 * std::sin(1.7 * n + 0.3));
  normal_model model(y);
  const stanx::model::model_base& base = model;
  std::vector<int> pi;

  // gradient through the tape vs the closed form
  double worst_g = 0.0, worst_lp = 0.0;
  for (int t = 0; t < 5; ++t) {
    std::vector<double> x = {1.0 + 0.7 * t, -0.5 + 0.3 * t}, g, gc;
    const double lp = stanx::model::log_prob_grad<true, true>(model, x, pi, g);
    const double lc = model.closed(x, gc);
    worst_lp = std::fmax(worst_lp, std::fabs(lp - lc) / std::fabs(lc));
    for (int i = 0; i < 2; ++i)
      worst_g = std::fmax(worst_g, std::fabs(g[i] - gc[i]) / (std::fabs(gc[i]) + 1e-300));
  }
  std::printf("crtp_lp_rel %.17g\ncrtp_grad_rel %.17g\n", worst_lp, worst_g);

  // the eight overloads through the base class
  {
    std::vector<double> x = {2.0, 0.4};
    const double a = base.log_prob(x, pi, nullptr);
    const double b = base.log_prob_jacobian(x, pi, nullptr);
    const double c = base.log_prob_propto(x, pi, nullptr);
    const double d = base.log_prob_propto_jacobian(x, pi, nullptr);
    std::printf("jacobian_term %.17g\n", b - a);  // = u
    std::printf("propto_jacobian_term %.17g\n", d - c);
    std::printf("constant_term %.17g\n", a - c);  // normalising constants
    double worst = 0.0;
    for (int k = 0; k < 4; ++k) {
      stanx::math::tape_scope scope;
      std::vector<var> v = {var(2.0), var(0.4)};
      const var lv = k == 0 ? base.log_prob(v, pi, nullptr)
                   : k == 1 ? base.log_prob_jacobian(v, pi, nullptr)
                   : k == 2 ? base.log_prob_propto(v, pi, nullptr)
                            : base.log_prob_propto_jacobian(v, pi, nullptr);
      const double ref = k == 0 ? a : k == 1 ? b : k == 2 ? c : d;
      worst = std::fmax(worst, std::fabs(lv.val() - ref));
    }
    std::printf("var_vs_double_maxdiff %.17g\n", worst);
  }

  // a precomputed-gradient entry inside a larger expression:
  // f(q) = 2 lp_gauss(q) + q0^2 = -|q|^2 + q0^2, df/dq = (0, -2 q1, -2 q2)
  {
    stanx::model::callback_model gm(gauss, nullptr, 3);
    stanx::math::tape_scope scope;
    std::vector<var> q = {var(0.5), var(-1.5), var(2.0)};
    std::vector<int> none;
    const var lp = gm.log_prob_propto_jacobian(q, none, nullptr);
    const var f = 2.0 * lp + stanx::math::square(q[0]);
    f.grad();
    const double err = std::fabs(q[0].adj() - 0.0) + std::fabs(q[1].adj() - 3.0) +
                       std::fabs(q[2].adj() + 4.0) + std::fabs(f.val() - (-6.5 + 0.25));
    std::printf("precomputed_node_err %.17g\n", err);
  }

  // an exception inside the tape path leaves the tape empty
  {
    std::vector<double> x = {0.0, 60.0}, g;
    bool threw = false;
    try {
      stanx::model::log_prob_grad<true, true>(model, x, pi, g);
    } catch (const std::domain_error&) {
      threw = true;
    }
    std::printf("threw %d\ntape_after_throw %zu\n", threw ? 1 : 0,
                stanx::math::active_tape().val.size());
  }

  // the adaptive service over the generated-style model
  {
    stanx::callbacks::interrupt interrupt;
    stanx::callbacks::logger logger;
    stanx::callbacks::writer init_w, diag_w;
    stanx::callbacks::table_writer sample_w;
    stanx::callbacks::structured_writer metric_w;
    const int rc = stanx::services::sample::hmc_nuts_diag_e_adapt(
        model, nullptr, nullptr, 4321u, 1u, 2.0, 1000, 4000, 1, false, 0, 1.0, 0.0, 10, 0.8,
        0.05, 0.75, 10.0, 75u, 50u, 25u, interrupt, logger, init_w, sample_w, diag_w, metric_w);
    double m = 0, m2 = 0;
    for (const auto& r : sample_w.rows) {
      m += r[7];
      m2 += r[7] * r[7];
    }
    const double n = static_cast<double>(sample_w.rows.size());
    m /= n;
    double ybar = 0;
    for (double v : y) ybar += v;
    ybar /= y.size();
    std::printf("service_rc %d\nservice_rows %.0f\nmu_mean %.17g\nmu_sd %.17g\nybar %.17g\n", rc,
                n, m, std::sqrt(m2 / n - m * m), ybar);
  }
  return 0;
}
