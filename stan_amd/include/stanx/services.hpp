// Sampler services, mirroring in arguments, call order and return codes:
//   error_codes            src/stan/services/error_codes.hpp:7-18
//   util::initialize       src/stan/services/util/initialize.hpp:68-227
//   util::generate_transitions
//                          src/stan/services/util/generate_transitions.hpp:42-77
//   util::mcmc_writer      src/stan/services/util/mcmc_writer.hpp:22-260
//   util::run_adaptive_sampler
//                          src/stan/services/util/run_adaptive_sampler.hpp:46-103
//   util::run_sampler      src/stan/services/util/run_sampler.hpp:39-85
//   sample::hmc_nuts_diag_e_adapt
//                          src/stan/services/sample/hmc_nuts_diag_e_adapt.hpp:57-112
//   sample::hmc_nuts_diag_e
//                          src/stan/services/sample/hmc_nuts_diag_e.hpp:49-91
//   sample::hmc_nuts_dense_e_adapt
//                          src/stan/services/sample/hmc_nuts_dense_e_adapt.hpp:57-112
//   sample::hmc_nuts_dense_e
//                          src/stan/services/sample/hmc_nuts_dense_e.hpp:48-93
//   util::read_dense_inv_metric, util::validate_dense_inv_metric,
//   util::create_unit_e_dense_inv_metric
//                          src/stan/services/util/read_dense_inv_metric.hpp:25-42,
//                          validate_dense_inv_metric.hpp:18-28,
//                          create_unit_e_dense_inv_metric.hpp:18-27
#ifndef STANX_SERVICES_HPP
#define STANX_SERVICES_HPP

#include <chrono>
#include <cmath>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "callbacks.hpp"
#include "io.hpp"
#include "mcmc/adapt_dense_e_nuts.hpp"
#include "mcmc/adapt_diag_e_nuts.hpp"
#include "model.hpp"
#include "philox.hpp"

namespace stanx {
namespace services {

struct error_codes {
  enum { OK = 0, USAGE = 64, DATAERR = 65, NOINPUT = 66, SOFTWARE = 70, CONFIG = 78 };
};

namespace util {

inline rng_t create_rng(unsigned int seed, unsigned int chain) {
  return rng_t(seed, chain);
}

// Random (or user-supplied) unconstrained inits, checked for a finite log
// density and gradient; up to 100 tries.  Inits are uniform(-R, R) per
// coordinate (io/random_var_context.hpp:42-66).  n_evals (optional) counts
// the gradient evaluations made.
template <class Model>
std::vector<double> initialize(const Model& model, const std::vector<double>* init,
                               rng_t& rng, double init_radius, bool print_timing,
                               callbacks::logger& logger,
                               callbacks::writer& init_writer,
                               long long* n_evals = nullptr) {
  const size_t D = model.num_params_r();
  std::vector<double> unconstrained(D);
  const bool zero_init = init_radius == 0.0;
  const int max_tries = (init || zero_init) ? 1 : 100;
  std::vector<int> params_i;
  for (int t = 0; t < max_tries; ++t) {
    if (init) {
      unconstrained = *init;
    } else {
      for (size_t i = 0; i < D; ++i)
        unconstrained[i] =
            zero_init ? 0.0 : -init_radius + rng.uniform() * 2.0 * init_radius;
    }
    std::vector<double> gradient;
    double lp = 0;
    auto start = std::chrono::steady_clock::now();
    if (n_evals) ++*n_evals;
    try {
      lp = stanx::model::log_prob_grad<true, true>(model, unconstrained,
                                                   params_i, gradient);
    } catch (const std::domain_error& e) {
      logger.warn("Rejecting initial value:");
      logger.warn("  Error evaluating the log probability at the initial value.");
      logger.warn(e.what());
      continue;
    }
    auto end = std::chrono::steady_clock::now();
    if (!std::isfinite(lp)) {
      logger.warn("Rejecting initial value:");
      logger.warn("  Log probability evaluates to log(0), i.e. negative infinity.");
      continue;
    }
    double gsum = 0.0;
    for (double g : gradient) gsum += g;
    if (!std::isfinite(gsum)) {
      logger.warn("Rejecting initial value:");
      logger.warn("  Gradient evaluated at the initial value is not finite.");
      continue;
    }
    if (print_timing) {
      const double dt =
          std::chrono::duration_cast<std::chrono::microseconds>(end - start).count() /
          1e6;
      std::stringstream m1;
      m1 << "Gradient evaluation took " << dt << " seconds";
      logger.info(m1);
    }
    init_writer(unconstrained);
    return unconstrained;
  }
  if (!zero_init) {
    std::stringstream msg;
    msg << "Initialization between (-" << init_radius << ", " << init_radius
        << ") failed after " << max_tries << " attempts. ";
    logger.error(msg);
  }
  throw std::domain_error("Initialization failed.");
}

// sample writer: lp__, accept_stat__, sampler params, constrained params
template <class Sampler, class Model>
class mcmc_writer {
 public:
  mcmc_writer(callbacks::writer& sample_writer, callbacks::writer& diagnostic_writer,
              callbacks::logger& logger)
      : sample_writer_(sample_writer), diagnostic_writer_(diagnostic_writer),
        logger_(logger) {}
  void write_sample_names(Sampler& sampler, const Model& model) {
    std::vector<std::string> names;
    mcmc::sample::get_sample_param_names(names);
    sampler.get_sampler_param_names(names);
    model.constrained_param_names(names, true, true);
    sample_writer_(names);
  }
  void write_sample_params(rng_t& rng, const mcmc::sample& s, Sampler& sampler,
                           const Model& model) {
    std::vector<double> values;
    s.get_sample_params(values);
    sampler.get_sampler_params(values);
    std::vector<double> cont(s.cont_params()), model_values;
    std::vector<int> params_i;
    model.write_array(rng, cont, params_i, model_values, true, true, nullptr);
    values.insert(values.end(), model_values.begin(), model_values.end());
    sample_writer_(values);
  }
  // unconstrained parameters, then their momenta and gradients
  // (mcmc_writer.hpp:157-190)
  void write_diagnostic_names(Sampler& sampler, const Model& model) {
    std::vector<std::string> names, model_names;
    mcmc::sample::get_sample_param_names(names);
    sampler.get_sampler_param_names(names);
    model.unconstrained_param_names(model_names, false, false);
    sampler.get_sampler_diagnostic_names(model_names, names);
    diagnostic_writer_(names);
  }
  void write_diagnostic_params(const mcmc::sample& s, Sampler& sampler) {
    std::vector<double> values;
    s.get_sample_params(values);
    sampler.get_sampler_params(values);
    sampler.get_sampler_diagnostics(values);
    diagnostic_writer_(values);
  }
  void write_adapt_finish(Sampler& sampler) {
    sample_writer_("Adaptation terminated");
    std::stringstream ss;
    ss << "Step size = " << sampler.get_nominal_stepsize();
    sample_writer_(ss.str());
    sampler.z().write_metric(sample_writer_);
  }
  // blank line, three aligned lines, blank line (mcmc_writer.hpp:203-238)
  void write_timing(double warm, double samp, callbacks::writer& w) {
    const std::string title(" Elapsed Time: ");
    w();
    std::stringstream s1, s2, s3;
    s1 << title << warm << " seconds (Warm-up)";
    w(s1.str());
    s2 << std::string(title.size(), ' ') << samp << " seconds (Sampling)";
    w(s2.str());
    s3 << std::string(title.size(), ' ') << warm + samp << " seconds (Total)";
    w(s3.str());
    w();
  }
  void write_timing(double warm, double samp) {
    write_timing(warm, samp, sample_writer_);
    write_timing(warm, samp, diagnostic_writer_);
  }

 private:
  callbacks::writer& sample_writer_;
  callbacks::writer& diagnostic_writer_;
  callbacks::logger& logger_;
};

template <class Sampler, class Model>
void generate_transitions(Sampler& sampler, int num_iterations, int start, int finish,
                          int num_thin, int refresh, bool save, bool warmup,
                          mcmc_writer<Sampler, Model>& writer, mcmc::sample& init_s,
                          const Model& model, rng_t& base_rng,
                          callbacks::interrupt& callback, callbacks::logger& logger,
                          size_t chain_id = 1, size_t num_chains = 1) {
  for (int m = 0; m < num_iterations; ++m) {
    callback();
    if (refresh > 0 && (start + m + 1 == finish || m == 0 || (m + 1) % refresh == 0)) {
      int w = static_cast<int>(std::ceil(std::log10(static_cast<double>(finish))));
      std::stringstream message;
      if (num_chains != 1) message << "Chain [" << chain_id << "] ";
      message << "Iteration: " << std::setw(w) << m + 1 + start << " / " << finish
              << " [" << std::setw(3)
              << static_cast<int>((100.0 * (start + m + 1)) / finish) << "%] "
              << (warmup ? " (Warmup)" : " (Sampling)");
      logger.info(message);
    }
    init_s = sampler.transition(init_s, logger);
    if (save && ((m % num_thin) == 0)) {
      writer.write_sample_params(base_rng, init_s, sampler, model);
      writer.write_diagnostic_params(init_s, sampler);
    }
  }
}

template <class Sampler, class Model>
void run_adaptive_sampler(Sampler& sampler, const Model& model,
                          std::vector<double>& cont_vector, int num_warmup,
                          int num_samples, int num_thin, int refresh, bool save_warmup,
                          rng_t& rng, callbacks::interrupt& interrupt,
                          callbacks::logger& logger, callbacks::writer& sample_writer,
                          callbacks::writer& diagnostic_writer,
                          callbacks::structured_writer& metric_writer, double* warm_s = nullptr,
                          double* samp_s = nullptr, size_t chain_id = 1,
                          size_t num_chains = 1) {
  sampler.engage_adaptation();
  try {
    sampler.z().q = cont_vector;
    sampler.init_stepsize(logger);
  } catch (const std::exception& e) {
    logger.error("Exception initializing step size.");
    logger.error(e.what());
    return;
  }
  mcmc_writer<Sampler, Model> writer(sample_writer, diagnostic_writer, logger);
  mcmc::sample s(cont_vector, 0, 0);
  writer.write_sample_names(sampler, model);
  writer.write_diagnostic_names(sampler, model);
  auto t0 = std::chrono::steady_clock::now();
  generate_transitions(sampler, num_warmup, 0, num_warmup + num_samples, num_thin,
                       refresh, save_warmup, true, writer, s, model, rng, interrupt,
                       logger, chain_id, num_chains);
  auto t1 = std::chrono::steady_clock::now();
  sampler.disengage_adaptation();
  writer.write_adapt_finish(sampler);
  sampler.write_sampler_state_struct(metric_writer);
  auto t2 = std::chrono::steady_clock::now();
  generate_transitions(sampler, num_samples, num_warmup, num_warmup + num_samples,
                       num_thin, refresh, true, false, writer, s, model, rng, interrupt,
                       logger, chain_id, num_chains);
  auto t3 = std::chrono::steady_clock::now();
  const double w = std::chrono::duration<double>(t1 - t0).count();
  const double sm = std::chrono::duration<double>(t3 - t2).count();
  writer.write_timing(w, sm);
  if (warm_s) *warm_s = w;
  if (samp_s) *samp_s = sm;
}

// Warm-up and sampling without adaptation; unlike run_adaptive_sampler no
// step-size heuristic runs first.
template <class Sampler, class Model>
void run_sampler(Sampler& sampler, const Model& model, std::vector<double>& cont_vector,
                 int num_warmup, int num_samples, int num_thin, int refresh,
                 bool save_warmup, rng_t& rng, callbacks::interrupt& interrupt,
                 callbacks::logger& logger, callbacks::writer& sample_writer,
                 callbacks::writer& diagnostic_writer, double* warm_s = nullptr,
                 double* samp_s = nullptr, size_t chain_id = 1, size_t num_chains = 1) {
  mcmc_writer<Sampler, Model> writer(sample_writer, diagnostic_writer, logger);
  mcmc::sample s(cont_vector, 0, 0);
  sampler.z().q = cont_vector;
  writer.write_sample_names(sampler, model);
  writer.write_diagnostic_names(sampler, model);
  auto t0 = std::chrono::steady_clock::now();
  generate_transitions(sampler, num_warmup, 0, num_warmup + num_samples, num_thin,
                       refresh, save_warmup, true, writer, s, model, rng, interrupt,
                       logger, chain_id, num_chains);
  auto t1 = std::chrono::steady_clock::now();
  writer.write_adapt_finish(sampler);
  auto t2 = std::chrono::steady_clock::now();
  generate_transitions(sampler, num_samples, num_warmup, num_warmup + num_samples,
                       num_thin, refresh, true, false, writer, s, model, rng, interrupt,
                       logger, chain_id, num_chains);
  auto t3 = std::chrono::steady_clock::now();
  const double w = std::chrono::duration<double>(t1 - t0).count();
  const double sm = std::chrono::duration<double>(t3 - t2).count();
  writer.write_timing(w, sm);
  if (warm_s) *warm_s = w;
  if (samp_s) *samp_s = sm;
}

}  // namespace util

namespace sample {

// hmc_nuts_diag_e: the same transition, no adaptation.
template <class Model>
int hmc_nuts_diag_e(const Model& model, const std::vector<double>* init,
                    const std::vector<double>* init_inv_metric, unsigned int random_seed,
                    unsigned int chain, double init_radius, int num_warmup,
                    int num_samples, int num_thin, bool save_warmup, int refresh,
                    double stepsize, double stepsize_jitter, int max_depth,
                    callbacks::interrupt& interrupt, callbacks::logger& logger,
                    callbacks::writer& init_writer, callbacks::writer& sample_writer,
                    callbacks::writer& diagnostic_writer, long long* n_grad_out = nullptr,
                    double* warm_s = nullptr, double* samp_s = nullptr) {
  rng_t rng = util::create_rng(random_seed, chain);
  std::vector<double> cont_vector;
  std::vector<double> inv_metric;
  long long init_evals = 0;
  try {
    cont_vector = util::initialize(model, init, rng, init_radius, true, logger,
                                   init_writer, &init_evals);
    if (init_inv_metric) {
      inv_metric = *init_inv_metric;
      if (inv_metric.size() != model.num_params_r())
        throw std::domain_error("inverse metric has the wrong size");
    } else {
      inv_metric.assign(model.num_params_r(), 1.0);
    }
    for (double v : inv_metric)
      if (!(v > 0) || !std::isfinite(v))
        throw std::domain_error("Inverse metric must be positive and finite");
  } catch (const std::exception& e) {
    logger.error(e.what());
    return error_codes::CONFIG;
  }
  mcmc::diag_e_nuts<Model> sampler(model, rng);
  sampler.set_metric(inv_metric);
  sampler.set_nominal_stepsize(stepsize);
  sampler.set_stepsize_jitter(stepsize_jitter);
  sampler.set_max_depth(max_depth);
  int rc = error_codes::OK;
  try {
    util::run_sampler(sampler, model, cont_vector, num_warmup, num_samples, num_thin,
                      refresh, save_warmup, rng, interrupt, logger, sample_writer,
                      diagnostic_writer, warm_s, samp_s);
  } catch (const std::exception& e) {
    logger.error(e.what());
    rc = error_codes::SOFTWARE;
  }
  if (n_grad_out) *n_grad_out = init_evals + sampler.hamiltonian().n_grad();
  return rc;
}

// hmc_nuts_diag_e_adapt with explicit init / metric vectors (the
// var_context forms of the reference reduce to these after reading).
template <class Model>
int hmc_nuts_diag_e_adapt(const Model& model, const std::vector<double>* init,
                          const std::vector<double>* init_inv_metric,
                          unsigned int random_seed, unsigned int chain,
                          double init_radius, int num_warmup, int num_samples,
                          int num_thin, bool save_warmup, int refresh, double stepsize,
                          double stepsize_jitter, int max_depth, double delta,
                          double gamma, double kappa, double t0,
                          unsigned int init_buffer, unsigned int term_buffer,
                          unsigned int window, callbacks::interrupt& interrupt,
                          callbacks::logger& logger, callbacks::writer& init_writer,
                          callbacks::writer& sample_writer,
                          callbacks::writer& diagnostic_writer,
                          callbacks::structured_writer& metric_writer,
                          double* stepsize_out = nullptr,
                          std::vector<double>* inv_metric_out = nullptr,
                          long long* n_grad_out = nullptr,
                          double* warm_s = nullptr, double* samp_s = nullptr) {
  rng_t rng = util::create_rng(random_seed, chain);
  std::vector<double> cont_vector;
  std::vector<double> inv_metric;
  long long init_evals = 0;
  try {
    cont_vector = util::initialize(model, init, rng, init_radius, true, logger,
                                   init_writer, &init_evals);
    if (init_inv_metric) {
      inv_metric = *init_inv_metric;
      if (inv_metric.size() != model.num_params_r())
        throw std::domain_error("inverse metric has the wrong size");
    } else {
      inv_metric.assign(model.num_params_r(), 1.0);
    }
    for (double v : inv_metric)
      if (!(v > 0) || !std::isfinite(v))
        throw std::domain_error("Inverse metric must be positive and finite");
  } catch (const std::exception& e) {
    logger.error(e.what());
    return error_codes::CONFIG;
  }
  mcmc::adapt_diag_e_nuts<Model> sampler_obj(model, rng);
  auto* sampler = &sampler_obj;
  sampler->set_metric(inv_metric);
  sampler->set_nominal_stepsize(stepsize);
  sampler->set_stepsize_jitter(stepsize_jitter);
  sampler->set_max_depth(max_depth);
  sampler->get_stepsize_adaptation().set_mu(std::log(10 * stepsize));
  sampler->get_stepsize_adaptation().set_delta(delta);
  sampler->get_stepsize_adaptation().set_gamma(gamma);
  sampler->get_stepsize_adaptation().set_kappa(kappa);
  sampler->get_stepsize_adaptation().set_t0(t0);
  sampler->set_window_params(num_warmup, init_buffer, term_buffer, window, logger);
  int rc = error_codes::OK;
  try {
    util::run_adaptive_sampler(*sampler, model, cont_vector, num_warmup, num_samples,
                               num_thin, refresh, save_warmup, rng, interrupt, logger,
                               sample_writer, diagnostic_writer, metric_writer, warm_s,
                               samp_s);
  } catch (const std::exception& e) {
    logger.error(e.what());
    rc = error_codes::SOFTWARE;
  }
  if (stepsize_out) *stepsize_out = sampler->get_nominal_stepsize();
  if (inv_metric_out) *inv_metric_out = sampler->z().inv_e_metric_;
  if (n_grad_out) *n_grad_out = init_evals + sampler->hamiltonian().n_grad();
  return rc;
}

// The dense inverse metric (row-major D x D): the given one, or the
// identity (create_unit_e_dense_inv_metric); it must be D x D, finite,
// symmetric to 1e-8 and positive definite (validate_dense_inv_metric:
// "Inverse Euclidean metric not positive definite." and an
// "Initialization failure" domain error otherwise).
inline std::vector<double> dense_inv_metric(const std::vector<double>* given, size_t D,
                                            callbacks::logger& logger) {
  std::vector<double> m(D * D, 0.0);
  if (given) {
    if (given->size() != D * D) {
      logger.error("Cannot get inverse metric from input file.");
      throw std::domain_error("Initialization failure");
    }
    m = *given;
  } else {
    for (size_t i = 0; i < D; ++i) m[i * D + i] = 1.0;
  }
  bool ok = true;
  for (size_t i = 0; i < D && ok; ++i)
    for (size_t j = 0; j < D; ++j)
      if (!std::isfinite(m[i * D + j]) || std::fabs(m[i * D + j] - m[j * D + i]) > 1e-8) {
        ok = false;
        break;
      }
  std::vector<double> L;
  if (!ok || !mcmc::cholesky_lower(m, static_cast<int>(D), L)) {
    logger.error("Inverse Euclidean metric not positive definite.");
    throw std::domain_error("Initialization failure");
  }
  return m;
}

// hmc_nuts_dense_e_adapt: windowed covariance adaptation of a dense
// metric; inv_metric_out receives the adapted D x D matrix.
template <class Model>
int hmc_nuts_dense_e_adapt(const Model& model, const std::vector<double>* init,
                           const std::vector<double>* init_inv_metric,
                           unsigned int random_seed, unsigned int chain,
                           double init_radius, int num_warmup, int num_samples,
                           int num_thin, bool save_warmup, int refresh, double stepsize,
                           double stepsize_jitter, int max_depth, double delta,
                           double gamma, double kappa, double t0,
                           unsigned int init_buffer, unsigned int term_buffer,
                           unsigned int window, callbacks::interrupt& interrupt,
                           callbacks::logger& logger, callbacks::writer& init_writer,
                           callbacks::writer& sample_writer,
                           callbacks::writer& diagnostic_writer,
                           callbacks::structured_writer& metric_writer,
                           double* stepsize_out = nullptr,
                           std::vector<double>* inv_metric_out = nullptr,
                           long long* n_grad_out = nullptr,
                           double* warm_s = nullptr, double* samp_s = nullptr) {
  rng_t rng = util::create_rng(random_seed, chain);
  std::vector<double> cont_vector;
  std::vector<double> inv_metric;
  long long init_evals = 0;
  try {
    cont_vector = util::initialize(model, init, rng, init_radius, true, logger,
                                   init_writer, &init_evals);
    inv_metric = dense_inv_metric(init_inv_metric, model.num_params_r(), logger);
  } catch (const std::exception& e) {
    logger.error(e.what());
    return error_codes::CONFIG;
  }
  mcmc::adapt_dense_e_nuts<Model> sampler(model, rng);
  sampler.set_metric(inv_metric);
  sampler.set_nominal_stepsize(stepsize);
  sampler.set_stepsize_jitter(stepsize_jitter);
  sampler.set_max_depth(max_depth);
  sampler.get_stepsize_adaptation().set_mu(std::log(10 * stepsize));
  sampler.get_stepsize_adaptation().set_delta(delta);
  sampler.get_stepsize_adaptation().set_gamma(gamma);
  sampler.get_stepsize_adaptation().set_kappa(kappa);
  sampler.get_stepsize_adaptation().set_t0(t0);
  sampler.set_window_params(num_warmup, init_buffer, term_buffer, window, logger);
  int rc = error_codes::OK;
  try {
    util::run_adaptive_sampler(sampler, model, cont_vector, num_warmup, num_samples,
                               num_thin, refresh, save_warmup, rng, interrupt, logger,
                               sample_writer, diagnostic_writer, metric_writer, warm_s,
                               samp_s);
  } catch (const std::exception& e) {
    logger.error(e.what());
    rc = error_codes::SOFTWARE;
  }
  if (stepsize_out) *stepsize_out = sampler.get_nominal_stepsize();
  if (inv_metric_out) *inv_metric_out = sampler.z().inv_e_metric_;
  if (n_grad_out) *n_grad_out = init_evals + sampler.hamiltonian().n_grad();
  return rc;
}

// hmc_nuts_dense_e: the dense-metric transition, no adaptation.
template <class Model>
int hmc_nuts_dense_e(const Model& model, const std::vector<double>* init,
                     const std::vector<double>* init_inv_metric, unsigned int random_seed,
                     unsigned int chain, double init_radius, int num_warmup,
                     int num_samples, int num_thin, bool save_warmup, int refresh,
                     double stepsize, double stepsize_jitter, int max_depth,
                     callbacks::interrupt& interrupt, callbacks::logger& logger,
                     callbacks::writer& init_writer, callbacks::writer& sample_writer,
                     callbacks::writer& diagnostic_writer, long long* n_grad_out = nullptr,
                     double* warm_s = nullptr, double* samp_s = nullptr) {
  rng_t rng = util::create_rng(random_seed, chain);
  std::vector<double> cont_vector;
  std::vector<double> inv_metric;
  long long init_evals = 0;
  try {
    cont_vector = util::initialize(model, init, rng, init_radius, true, logger,
                                   init_writer, &init_evals);
    inv_metric = dense_inv_metric(init_inv_metric, model.num_params_r(), logger);
  } catch (const std::exception& e) {
    logger.error(e.what());
    return error_codes::CONFIG;
  }
  mcmc::dense_e_nuts<Model> sampler(model, rng);
  sampler.set_metric(inv_metric);
  sampler.set_nominal_stepsize(stepsize);
  sampler.set_stepsize_jitter(stepsize_jitter);
  sampler.set_max_depth(max_depth);
  int rc = error_codes::OK;
  try {
    util::run_sampler(sampler, model, cont_vector, num_warmup, num_samples, num_thin,
                      refresh, save_warmup, rng, interrupt, logger, sample_writer,
                      diagnostic_writer, warm_s, samp_s);
  } catch (const std::exception& e) {
    logger.error(e.what());
    rc = error_codes::SOFTWARE;
  }
  if (n_grad_out) *n_grad_out = init_evals + sampler.hamiltonian().n_grad();
  return rc;
}

}  // namespace sample
}  // namespace services
}  // namespace stanx
#endif
