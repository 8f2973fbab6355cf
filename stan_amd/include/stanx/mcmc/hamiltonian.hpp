// Phase-space point and diagonal Euclidean Hamiltonian, mirroring
//   ps_point        src/stan/mcmc/hmc/hamiltonians/ps_point.hpp:16-52
//   diag_e_point    src/stan/mcmc/hmc/hamiltonians/diag_e_point.hpp:13-55
//   base_hamiltonian::update_potential_gradient
//                   src/stan/mcmc/hmc/hamiltonians/base_hamiltonian.hpp:61-70
//   diag_e_metric   src/stan/mcmc/hmc/hamiltonians/diag_e_metric.hpp:20-50
#ifndef STANX_MCMC_HAMILTONIAN_HPP
#define STANX_MCMC_HAMILTONIAN_HPP

#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../callbacks.hpp"
#include "../model.hpp"
#include "../philox.hpp"

namespace stanx {
namespace mcmc {

using vec_t = std::vector<double>;

class ps_point {
 public:
  explicit ps_point(int n) : q(n, 0.0), p(n, 0.0), g(n, 0.0) {}
  vec_t q, p, g;
  double V{0};
  // ps_point::operator= copies q, p, g and V (never the metric)
  void assign(const ps_point& o) {
    q = o.q;
    p = o.p;
    g = o.g;
    V = o.V;
  }
};

class diag_e_point : public ps_point {
 public:
  explicit diag_e_point(int n) : ps_point(n), inv_e_metric_(n, 1.0) {}
  vec_t inv_e_metric_;
  void set_metric(const vec_t& m) { inv_e_metric_ = m; }
};

template <class Model>
class diag_e_metric {
 public:
  explicit diag_e_metric(const Model& model) : model_(model) {}

  double T(const diag_e_point& z) const {
    double t = 0.0;
    for (size_t i = 0; i < z.p.size(); ++i)
      t += z.p[i] * (z.inv_e_metric_[i] * z.p[i]);
    return 0.5 * t;
  }
  double V(const diag_e_point& z) const { return z.V; }
  double H(const diag_e_point& z) const {
    double t = 0.0;
    for (size_t i = 0; i < z.p.size(); ++i)
      t += z.p[i] * (z.inv_e_metric_[i] * z.p[i]);
    return 0.5 * t + z.V;
  }
  void dtau_dp(const diag_e_point& z, vec_t& out) const {
    out.resize(z.p.size());
    for (size_t i = 0; i < z.p.size(); ++i) out[i] = z.inv_e_metric_[i] * z.p[i];
  }
  void sample_p(diag_e_point& z, rng_t& rng) const {
    for (size_t i = 0; i < z.p.size(); ++i)
      z.p[i] = rng.normal() / std::sqrt(z.inv_e_metric_[i]);
  }
  void init(diag_e_point& z, callbacks::logger& logger) {
    update_potential_gradient(z, logger);
  }
  // V = -lp, g = -grad lp; a domain error rejects the point (V = +inf)
  void update_potential_gradient(diag_e_point& z, callbacks::logger& logger) {
    ++n_grad_;
    try {
      double lp = 0.0;
      stanx::model::gradient(model_, z.q, lp, gbuf_, logger);
      z.V = -lp;
      z.g.swap(gbuf_);
    } catch (const std::domain_error& e) {
      write_error_msg_(e, logger);
      z.V = std::numeric_limits<double>::infinity();
    }
    for (auto& gi : z.g) gi = -gi;
  }
  long long n_grad() const { return n_grad_; }
  void reset_n_grad() { n_grad_ = 0; }

 private:
  const Model& model_;
  vec_t gbuf_;
  long long n_grad_ = 0;
  void write_error_msg_(const std::exception& e, callbacks::logger& logger) {
    logger.error(
        "Informational Message: The current Metropolis proposal "
        "is about to be rejected because of the following issue:");
    logger.error(e.what());
  }
};

// expl_leapfrog (src/stan/mcmc/hmc/integrators/expl_leapfrog.hpp:16-32)
// driven by base_leapfrog::evolve (base_leapfrog.hpp:17-22)
template <class Hamiltonian>
class expl_leapfrog {
 public:
  void begin_update_p(diag_e_point& z, Hamiltonian& h, double epsilon,
                      callbacks::logger&) {
    for (size_t i = 0; i < z.p.size(); ++i) z.p[i] -= epsilon * z.g[i];
  }
  void update_q(diag_e_point& z, Hamiltonian& h, double epsilon,
                callbacks::logger& logger) {
    for (size_t i = 0; i < z.q.size(); ++i)
      z.q[i] += epsilon * (z.inv_e_metric_[i] * z.p[i]);
    h.update_potential_gradient(z, logger);
  }
  void end_update_p(diag_e_point& z, Hamiltonian& h, double epsilon,
                    callbacks::logger&) {
    for (size_t i = 0; i < z.p.size(); ++i) z.p[i] -= epsilon * z.g[i];
  }
  void evolve(diag_e_point& z, Hamiltonian& h, double epsilon,
              callbacks::logger& logger) {
    begin_update_p(z, h, 0.5 * epsilon, logger);
    update_q(z, h, epsilon, logger);
    end_update_p(z, h, 0.5 * epsilon, logger);
  }
};

}  // namespace mcmc
}  // namespace stanx
#endif
