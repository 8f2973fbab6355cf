// dense_e NUTS, plain and adaptive: the diag_e samplers with a dense
// inverse metric, adapted by windowed covariance estimation, mirroring
//   dense_e_nuts            src/stan/mcmc/hmc/nuts/dense_e_nuts.hpp:15-21
//   adapt_dense_e_nuts      src/stan/mcmc/hmc/nuts/adapt_dense_e_nuts.hpp:17-49
//   stepsize_covar_adapter  src/stan/mcmc/stepsize_covar_adapter.hpp:13-37
#ifndef STANX_MCMC_ADAPT_DENSE_E_NUTS_HPP
#define STANX_MCMC_ADAPT_DENSE_E_NUTS_HPP

#include <cmath>
#include <vector>

#include "adaptation.hpp"
#include "base_nuts.hpp"

namespace stanx {
namespace mcmc {

// NUTS with a dense metric and no adaptation (hmc_nuts_dense_e)
template <class Model>
class dense_e_nuts : public base_nuts<Model, dense_e_metric<Model>> {
 public:
  dense_e_nuts(const Model& model, rng_t& rng)
      : base_nuts<Model, dense_e_metric<Model>>(model, rng) {}
};

template <class Model>
class adapt_dense_e_nuts : public dense_e_nuts<Model> {
 public:
  adapt_dense_e_nuts(const Model& model, rng_t& rng)
      : dense_e_nuts<Model>(model, rng),
        covar_adaptation_(static_cast<int>(model.num_params_r()), "covariance") {}

  void engage_adaptation() { adapt_flag_ = true; }
  void disengage_adaptation() {
    adapt_flag_ = false;
    this->nom_epsilon_ = stepsize_adaptation_.final_step_size();
  }
  bool adapting() const { return adapt_flag_; }
  step_size_tuner& get_stepsize_adaptation() { return stepsize_adaptation_; }
  covar_adaptation& get_covar_adaptation() { return covar_adaptation_; }
  void set_window_params(unsigned int num_warmup, unsigned int init_buffer,
                         unsigned int term_buffer, unsigned int base_window,
                         callbacks::logger& logger) {
    covar_adaptation_.set_window_params(num_warmup, init_buffer, term_buffer,
                                        base_window, logger);
  }

  // the transition, then dual averaging and the covariance window; a new
  // metric brings a new Cholesky factor and a fresh step-size heuristic
  sample transition(const sample& init_sample, callbacks::logger& logger) {
    sample s = base_nuts<Model, dense_e_metric<Model>>::transition(init_sample, logger);
    if (adapt_flag_) {
      this->nom_epsilon_ = stepsize_adaptation_.update(s.accept_stat());
      const bool update =
          covar_adaptation_.observe(this->z_.inv_e_metric_, this->z_.q);
      if (update) {
        this->z_.factorize();
        this->init_stepsize(logger);
        stepsize_adaptation_.set_mu(std::log(10 * this->nom_epsilon_));
        stepsize_adaptation_.restart();
      }
    }
    return s;
  }

 protected:
  bool adapt_flag_ = false;
  step_size_tuner stepsize_adaptation_;
  covar_adaptation covar_adaptation_;
};

}  // namespace mcmc
}  // namespace stanx
#endif
