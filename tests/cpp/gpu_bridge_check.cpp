// The GPU log density inside a model written in the generated style: the
// GLM (fused HIP kernels through the C-ABI) enters the model's tape as one
// precomputed-gradient entry, and the model adds a term of its own on the
// same tape (a ridge penalty -lambda/2 |beta|^2).  Checks, printed as
// "name value" lines for the pytest side:
//   - value and gradient of the composite against sx_glm_logp_grad plus the
//     closed form of the extra term;
//   - the adaptive service over the composite model runs (rc, rows).
// Linked against stan_amd/libstanx.so; needs a GPU.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/stanx.h"
#include "../../stan_amd/include/stanx/callbacks.hpp"
#include "../../stan_amd/include/stanx/gpu_glm_model.hpp"
#include "../../stan_amd/include/stanx/model.hpp"
#include "../../stan_amd/include/stanx/services.hpp"

using stanx::math::var;

class ridge_glm : public stanx::model::model_base_crtp<ridge_glm> {
 public:
  ridge_glm(const stanx::model::gpu_glm_model& glm, double lambda)
      : model_base_crtp(glm.num_params_r()), glm_(glm), lambda_(lambda) {}

  template <bool propto, bool jacobian, typename T>
  T log_prob(std::vector<T>& params_r, std::vector<int>& params_i,
             std::ostream* msgs = nullptr) const {
    // the GLM with its priors: one entry (T = var) or one value (T = double)
    T lp = glm_.template log_prob<propto, jacobian>(params_r, params_i, msgs);
    for (size_t k = 1; k < params_r.size(); ++k)
      lp += -0.5 * lambda_ * params_r[k] * params_r[k];
    return lp;
  }

  std::string model_name() const override { return "ridge_glm"; }
  void get_param_names(std::vector<std::string>& n, bool a, bool b) const override {
    glm_.get_param_names(n, a, b);
  }
  void get_dims(std::vector<std::vector<size_t>>& d, bool a, bool b) const override {
    glm_.get_dims(d, a, b);
  }
  void constrained_param_names(std::vector<std::string>& n, bool a, bool b) const override {
    glm_.constrained_param_names(n, a, b);
  }
  void unconstrained_param_names(std::vector<std::string>& n, bool a, bool b) const override {
    glm_.unconstrained_param_names(n, a, b);
  }
  void transform_inits(const stanx::io::var_context& c, std::vector<int>& pi,
                       std::vector<double>& r, std::ostream* m) const override {
    glm_.transform_inits(c, pi, r, m);
  }
  void write_array(stanx::rng_t& rng, std::vector<double>& r, std::vector<int>& pi,
                   std::vector<double>& vars, bool a, bool b, std::ostream* m) const override {
    glm_.write_array(rng, r, pi, vars, a, b, m);
  }
  void unconstrain_array(const std::vector<double>& c, std::vector<double>& u,
                         std::ostream* m) const override {
    glm_.unconstrain_array(c, u, m);
  }

 private:
  const stanx::model::gpu_glm_model& glm_;
  double lambda_;
};

int main() {
  const int64_t N = 3000, K = 6;
  std::vector<double> X(N * K);
  std::vector<int32_t> y(N);
  for (int64_t n = 0; n < N; ++n) {
    double eta = 0.3;
    for (int64_t k = 0; k < K; ++k) {
      X[n * K + k] = 0.4 * std::sin(0.37 * n + 1.3 * k) + 0.2 * std::cos(0.11 * n * (k + 1));
      eta += X[n * K + k] * (0.5 - 0.2 * k);
    }
    y[n] = std::sin(12.9898 * n) * 0.5 + 0.5 < 1.0 / (1.0 + std::exp(-eta)) ? 1 : 0;
  }
  sx_prior prior{5.0, 2.0, 2.0, 1.0, 0.0};
  sx_glm* ctx = nullptr;
  if (sx_glm_create(&ctx, SX_LOGIT, X.data(), y.data(), N, K, nullptr, 0, SX_FP64, &prior, 0) !=
      SX_OK) {
    std::printf("create_failed 1\n");
    std::fprintf(stderr, "%s\n", sx_last_error());
    return 1;
  }
  {
    stanx::model::gpu_glm_model glm(ctx);
    const double lambda = 0.7;
    ridge_glm model(glm, lambda);
    std::vector<int> pi;
    double worst_lp = 0.0, worst_g = 0.0;
    for (int t = 0; t < 4; ++t) {
      std::vector<double> theta(K + 1), g, gref(K + 1);
      for (int i = 0; i <= K; ++i) theta[i] = 0.3 * std::sin(1.1 * i + t) - 0.1 * t;
      const double lp = stanx::model::log_prob_grad<true, true>(model, theta, pi, g);
      double lref = 0.0;
      sx_glm_logp_grad(ctx, theta.data(), 1, 1, &lref, gref.data());
      for (int k = 1; k <= K; ++k) {
        lref += -0.5 * lambda * theta[k] * theta[k];
        gref[k] += -lambda * theta[k];
      }
      worst_lp = std::fmax(worst_lp, std::fabs(lp - lref) / std::fabs(lref));
      double num = 0.0, den = 0.0;
      for (int i = 0; i <= K; ++i) {
        num += (g[i] - gref[i]) * (g[i] - gref[i]);
        den += gref[i] * gref[i];
      }
      worst_g = std::fmax(worst_g, std::sqrt(num / den));
    }
    std::printf("bridge_lp_rel %.17g\nbridge_grad_rel %.17g\n", worst_lp, worst_g);

    stanx::callbacks::interrupt interrupt;
    stanx::callbacks::logger logger;
    stanx::callbacks::writer init_w, diag_w;
    stanx::callbacks::table_writer sample_w;
    stanx::callbacks::structured_writer metric_w;
    const int rc = stanx::services::sample::hmc_nuts_diag_e_adapt(
        model, nullptr, nullptr, 99u, 1u, 2.0, 200, 200, 1, false, 0, 1.0, 0.0, 10, 0.8, 0.05,
        0.75, 10.0, 75u, 50u, 25u, interrupt, logger, init_w, sample_w, diag_w, metric_w);
    std::printf("service_rc %d\nservice_rows %zu\n", rc, sample_w.rows.size());
  }
  sx_glm_destroy(ctx);
  return 0;
}
