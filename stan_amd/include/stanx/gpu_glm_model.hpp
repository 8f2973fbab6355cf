// gpu_glm_model: a model_base whose log density and gradient come from the
// fused HIP kernels through the C-ABI (include/stanx.h).  It is the drop-in
// the reference's services take as their Model template parameter
// (src/stan/services/sample/hmc_nuts_diag_e_adapt.hpp:57-58) in place of a
// stanc3-generated model with
//   target += bernoulli_logit_glm_lpmf(y | X, alpha, beta)   (logistic)
//   target += normal_id_glm_lpdf(y | X, alpha, beta, sigma)  (normal)
// plus normal priors, and the hierarchical logistic model (varying
// intercepts and slopes by group, non-centred: z ~ normal(0, 1),
// tau ~ half-normal) whose specification analogue is
// src/test/test-models/good/variational/hier_logistic.stan.  Errors follow the reference: SX_EDOMAIN becomes
// std::domain_error (a rejected proposal), anything else std::runtime_error.
#ifndef STANX_GPU_GLM_MODEL_HPP
#define STANX_GPU_GLM_MODEL_HPP

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/stanx.h"
#include "model.hpp"

namespace stanx {
namespace model {

class gpu_glm_model : public model_base {
 public:
  // wraps a context created by sx_glm_create; does not take ownership
  explicit gpu_glm_model(sx_glm* ctx)
      : model_base(num_params_of(ctx)), ctx_(ctx) {
    family_ = infer_family();
    int64_t K = 0, J = 0;
    if (sx_glm_dims(ctx, nullptr, &K, &J) != SX_OK) throw std::runtime_error(sx_last_error());
    K_ = static_cast<size_t>(K);
    J_ = static_cast<size_t>(J);
  }

  sx_glm* context() const { return ctx_; }
  std::string model_name() const override {
    return family_ == SX_NORMAL   ? "normal_id_glm_gpu"
           : family_ == SX_LOGIT ? "bernoulli_logit_glm_gpu"
                                 : "hier_logistic_gpu";
  }
  void get_param_names(std::vector<std::string>& names, bool, bool) const override {
    names = {"alpha", "beta"};
    if (family_ == SX_NORMAL) names.push_back("sigma");
    if (family_ == SX_HIER_LOGIT) {
      names.push_back("tau");
      names.push_back("z");
    }
  }
  void get_dims(std::vector<std::vector<size_t>>& dimss, bool, bool) const override {
    dimss = {{}, {K()}};
    if (family_ == SX_NORMAL) dimss.push_back({});
    if (family_ == SX_HIER_LOGIT) {
      dimss.push_back({K() + 1});
      dimss.push_back({J_, K() + 1});
    }
  }
  // names in the order write_array emits the values (z group by group)
  void constrained_param_names(std::vector<std::string>& names, bool,
                               bool) const override {
    names.push_back("alpha");
    for (size_t k = 0; k < K(); ++k) names.push_back("beta." + std::to_string(k + 1));
    if (family_ == SX_NORMAL) names.push_back("sigma");
    if (family_ == SX_HIER_LOGIT) {
      for (size_t k = 0; k <= K(); ++k) names.push_back("tau." + std::to_string(k + 1));
      for (size_t j = 0; j < J_; ++j)
        for (size_t k = 0; k <= K(); ++k)
          names.push_back("z." + std::to_string(j + 1) + "." + std::to_string(k + 1));
    }
  }
  void unconstrained_param_names(std::vector<std::string>& names, bool,
                                 bool) const override {
    constrained_param_names(names, true, true);
  }
  double log_prob_grad_impl(const std::vector<double>& params_r,
                            std::vector<double>* gradient, bool propto, bool jacobian,
                            std::ostream*) const override {
    double lp = 0.0;
    int rc = gradient ? sx_glm_logp_grad(ctx_, params_r.data(), propto, jacobian, &lp,
                                         gradient->data())
                      : sx_glm_logp(ctx_, params_r.data(), propto, jacobian, &lp);
    if (rc == SX_EDOMAIN) throw std::domain_error(sx_last_error());
    if (rc != SX_OK) throw std::runtime_error(sx_last_error());
    return lp;
  }
  void transform_inits(const io::var_context& context, std::vector<int>&,
                       std::vector<double>& params_r, std::ostream*) const override {
    params_r.clear();
    params_r.push_back(context.vals_r("alpha")[0]);
    auto b = context.vals_r("beta");
    params_r.insert(params_r.end(), b.begin(), b.end());
    if (family_ == SX_NORMAL) {
      const double s = context.vals_r("sigma")[0];
      if (!(s > 0)) throw std::domain_error("sigma must be positive");
      params_r.push_back(std::log(s));
    }
    if (family_ == SX_HIER_LOGIT) {
      for (double t : context.vals_r("tau")) {
        if (!(t > 0)) throw std::domain_error("tau must be positive");
        params_r.push_back(std::log(t));
      }
      // z is stored row by row (group-major) in theta
      auto z = context.vals_r("z");
      for (size_t j = 0; j < J_; ++j)
        for (size_t k = 0; k <= K(); ++k) params_r.push_back(z[k * J_ + j]);
    }
  }
  void write_array(rng_t&, std::vector<double>& params_r, std::vector<int>&,
                   std::vector<double>& vars, bool, bool, std::ostream*) const override {
    vars = params_r;
    if (family_ == SX_NORMAL) vars.back() = std::exp(vars.back());  // sigma = exp(u)
    if (family_ == SX_HIER_LOGIT)
      for (size_t k = 0; k <= K(); ++k) vars[1 + K() + k] = std::exp(vars[1 + K() + k]);
  }
  void unconstrain_array(const std::vector<double>& c, std::vector<double>& u,
                         std::ostream*) const override {
    u = c;
    if (family_ == SX_NORMAL) u.back() = std::log(u.back());
    if (family_ == SX_HIER_LOGIT)
      for (size_t k = 0; k <= K(); ++k) u[1 + K() + k] = std::log(u[1 + K() + k]);
  }

 private:
  static size_t num_params_of(sx_glm* ctx) {
    int64_t D = 0;
    if (sx_glm_num_params(ctx, &D) != SX_OK) throw std::runtime_error(sx_last_error());
    return static_cast<size_t>(D);
  }
  int infer_family() const { return sx_glm_family(ctx_); }
  size_t K() const { return K_; }
  sx_glm* ctx_;
  size_t K_ = 0, J_ = 0;
  int family_ = SX_LOGIT;
};

}  // namespace model
}  // namespace stanx
#endif
