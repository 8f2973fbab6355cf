// diag_e NUTS, plain and adaptive.  The adaptive sampler follows a
// transition with dual averaging of the step size and windowed variance
// estimation of the inverse metric, mirroring
//   diag_e_nuts           src/stan/mcmc/hmc/nuts/diag_e_nuts.hpp:15-22
//   adapt_diag_e_nuts     src/stan/mcmc/hmc/nuts/adapt_diag_e_nuts.hpp:15-49
//   stepsize_var_adapter  src/stan/mcmc/stepsize_var_adapter.hpp:12-36
//   base_adapter          src/stan/mcmc/base_adapter.hpp:7-19
#ifndef STANX_MCMC_ADAPT_DIAG_E_NUTS_HPP
#define STANX_MCMC_ADAPT_DIAG_E_NUTS_HPP

#include <cmath>
#include <string>
#include <vector>

#include "adaptation.hpp"
#include "base_nuts.hpp"

namespace stanx {
namespace mcmc {

// NUTS with a diagonal metric and no adaptation (hmc_nuts_diag_e)
template <class Model>
class diag_e_nuts : public base_nuts<Model> {
 public:
  diag_e_nuts(const Model& model, rng_t& rng) : base_nuts<Model>(model, rng) {}
};

template <class Model>
class adapt_diag_e_nuts : public diag_e_nuts<Model> {
 public:
  adapt_diag_e_nuts(const Model& model, rng_t& rng)
      : diag_e_nuts<Model>(model, rng),
        var_adaptation_(static_cast<int>(model.num_params_r()), "variance") {}

  void engage_adaptation() { adapt_flag_ = true; }
  void disengage_adaptation() {
    adapt_flag_ = false;
    this->nom_epsilon_ = stepsize_adaptation_.final_step_size();
  }
  bool adapting() const { return adapt_flag_; }
  step_size_tuner& get_stepsize_adaptation() { return stepsize_adaptation_; }
  var_adaptation& get_var_adaptation() { return var_adaptation_; }
  void set_window_params(unsigned int num_warmup, unsigned int init_buffer,
                         unsigned int term_buffer, unsigned int base_window,
                         callbacks::logger& logger) {
    var_adaptation_.set_window_params(num_warmup, init_buffer, term_buffer,
                                      base_window, logger);
  }

  sample transition(const sample& init_sample, callbacks::logger& logger) {
    sample s = base_nuts<Model>::transition(init_sample, logger);
    if (adapt_flag_) {
      this->nom_epsilon_ = stepsize_adaptation_.update(s.accept_stat());
      const bool update =
          var_adaptation_.observe(this->z_.inv_e_metric_, this->z_.q);
      if (update) {
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
  var_adaptation var_adaptation_;
};

}  // namespace mcmc
}  // namespace stanx
#endif
