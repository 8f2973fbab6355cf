// Phase-space points and the diagonal and dense Euclidean Hamiltonians,
// mirroring
//   ps_point        src/stan/mcmc/hmc/hamiltonians/ps_point.hpp:16-52
//   diag_e_point    src/stan/mcmc/hmc/hamiltonians/diag_e_point.hpp:13-55
//   base_hamiltonian::update_potential_gradient
//                   src/stan/mcmc/hmc/hamiltonians/base_hamiltonian.hpp:61-70
//   diag_e_metric   src/stan/mcmc/hmc/hamiltonians/diag_e_metric.hpp:20-50
//   dense_e_point   src/stan/mcmc/hmc/hamiltonians/dense_e_point.hpp:13-58
//   dense_e_metric  src/stan/mcmc/hmc/hamiltonians/dense_e_metric.hpp:21-54
#ifndef STANX_MCMC_HAMILTONIAN_HPP
#define STANX_MCMC_HAMILTONIAN_HPP

#include <cmath>
#include <limits>
#include <sstream>
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
  static const char* metric_type() { return "diag_e"; }
  void write_inv_metric(callbacks::structured_writer& w) const { w.write("inv_metric", inv_e_metric_); }
  // diag_e_point::write_metric (diag_e_point.hpp:44-54)
  void write_metric(callbacks::writer& writer) const {
    writer("Diagonal elements of inverse mass matrix:");
    std::stringstream m;
    for (size_t i = 0; i < inv_e_metric_.size(); ++i) m << (i ? ", " : "") << inv_e_metric_[i];
    writer(m.str());
  }
};

// Lower Cholesky factor A = L L^T of a row-major symmetric matrix (L
// row-major; column by column, k ascending in each sum).  False when a
// pivot is not positive.
inline bool cholesky_lower(const vec_t& A, int D, vec_t& L) {
  L.assign(static_cast<size_t>(D) * D, 0.0);
  for (int j = 0; j < D; ++j) {
    double s = A[static_cast<size_t>(j) * D + j];
    for (int k = 0; k < j; ++k) s -= L[static_cast<size_t>(j) * D + k] * L[static_cast<size_t>(j) * D + k];
    if (!(s > 0.0)) return false;
    const double d = std::sqrt(s);
    L[static_cast<size_t>(j) * D + j] = d;
    for (int i = j + 1; i < D; ++i) {
      double t = A[static_cast<size_t>(i) * D + j];
      for (int k = 0; k < j; ++k) t -= L[static_cast<size_t>(i) * D + k] * L[static_cast<size_t>(j) * D + k];
      L[static_cast<size_t>(i) * D + j] = t / d;
    }
  }
  return true;
}

// Point with a dense inverse metric (D x D, row-major; symmetric) and the
// Cholesky factor that sample_p uses, kept in step with the metric.
class dense_e_point : public ps_point {
 public:
  explicit dense_e_point(int n) : ps_point(n), inv_e_metric_(static_cast<size_t>(n) * n, 0.0) {
    for (int i = 0; i < n; ++i) inv_e_metric_[static_cast<size_t>(i) * n + i] = 1.0;
    factorize();
  }
  vec_t inv_e_metric_;
  vec_t factor_;  // lower Cholesky factor of inv_e_metric_
  static const char* metric_type() { return "dense_e"; }
  void write_inv_metric(callbacks::structured_writer& w) const {
    w.write("inv_metric", inv_e_metric_, q.size(), q.size());
  }
  void set_metric(const vec_t& m) {
    inv_e_metric_ = m;
    factorize();
  }
  // refresh the factor after inv_e_metric_ changed in place
  void factorize() {
    const int D = static_cast<int>(q.size());
    if (!cholesky_lower(inv_e_metric_, D, factor_))
      throw std::runtime_error("inverse metric is not positive definite");
  }
  // dense_e_point::write_metric (dense_e_point.hpp:44-56): one line per row
  void write_metric(callbacks::writer& writer) const {
    const size_t D = q.size();
    writer("Elements of inverse mass matrix:");
    if (D == 0) writer(std::string());
    for (size_t i = 0; i < D; ++i) {
      std::stringstream m;
      for (size_t j = 0; j < D; ++j) m << (j ? ", " : "") << inv_e_metric_[i * D + j];
      writer(m.str());
    }
  }
};

// What the diag and dense metrics share: the potential and its gradient
// through the model (base_hamiltonian::update_potential_gradient).
template <class Model>
class base_hamiltonian {
 public:
  explicit base_hamiltonian(const Model& model) : model_(model) {}
  // V = -lp, g = -grad lp; a domain error rejects the point (V = +inf)
  void update_potential_gradient(ps_point& z, callbacks::logger& logger) {
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
  void init(ps_point& z, callbacks::logger& logger) { update_potential_gradient(z, logger); }
  double V(const ps_point& z) const { return z.V; }
  long long n_grad() const { return n_grad_; }
  void reset_n_grad() { n_grad_ = 0; }

 protected:
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

// dense_e_metric: T = 1/2 p . (M^-1 p), p_sharp = M^-1 p (row-major
// matrix-vector product, j ascending), p = U^-1 u with M^-1 = U^T U.
template <class Model>
class dense_e_metric : public base_hamiltonian<Model> {
 public:
  using point_t = dense_e_point;
  explicit dense_e_metric(const Model& model) : base_hamiltonian<Model>(model) {}

  void dtau_dp(const dense_e_point& z, vec_t& out) const {
    const size_t D = z.p.size();
    out.resize(D);
    for (size_t i = 0; i < D; ++i) {
      const double* row = z.inv_e_metric_.data() + i * D;
      double s = 0.0;
      for (size_t j = 0; j < D; ++j) s += row[j] * z.p[j];
      out[i] = s;
    }
  }
  double T(const dense_e_point& z) const {
    const size_t D = z.p.size();
    double t = 0.0;
    for (size_t i = 0; i < D; ++i) {
      const double* row = z.inv_e_metric_.data() + i * D;
      double s = 0.0;
      for (size_t j = 0; j < D; ++j) s += row[j] * z.p[j];
      t += z.p[i] * s;
    }
    return 0.5 * t;
  }
  double H(const dense_e_point& z) const { return T(z) + z.V; }
  // u_i ~ N(0, 1) in index order, then U p = u by back substitution
  // (U = L^T, j ascending in each row's sum)
  void sample_p(dense_e_point& z, rng_t& rng) const {
    const size_t D = z.p.size();
    u_.resize(D);
    for (size_t i = 0; i < D; ++i) u_[i] = rng.normal();
    const vec_t& L = z.factor_;
    for (size_t ii = D; ii-- > 0;) {
      double t = u_[ii];
      for (size_t j = ii + 1; j < D; ++j) t -= L[j * D + ii] * z.p[j];
      z.p[ii] = t / L[ii * D + ii];
    }
  }
  // expl_leapfrog::update_q: q += eps p_sharp
  void update_q(dense_e_point& z, double epsilon) const {
    dtau_dp(z, ps_);
    for (size_t i = 0; i < z.q.size(); ++i) z.q[i] += epsilon * ps_[i];
  }

 private:
  mutable vec_t u_, ps_;
};

template <class Model>
class diag_e_metric : public base_hamiltonian<Model> {
 public:
  using point_t = diag_e_point;
  explicit diag_e_metric(const Model& model) : base_hamiltonian<Model>(model) {}

  double T(const diag_e_point& z) const {
    double t = 0.0;
    for (size_t i = 0; i < z.p.size(); ++i)
      t += z.p[i] * (z.inv_e_metric_[i] * z.p[i]);
    return 0.5 * t;
  }
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
  // expl_leapfrog::update_q: q += eps (minv * p)
  void update_q(diag_e_point& z, double epsilon) const {
    for (size_t i = 0; i < z.q.size(); ++i)
      z.q[i] += epsilon * (z.inv_e_metric_[i] * z.p[i]);
  }
};

// expl_leapfrog (src/stan/mcmc/hmc/integrators/expl_leapfrog.hpp:16-32)
// driven by base_leapfrog::evolve (base_leapfrog.hpp:17-22)
template <class Hamiltonian>
class expl_leapfrog {
 public:
  using point_t = typename Hamiltonian::point_t;
  void begin_update_p(point_t& z, Hamiltonian& h, double epsilon,
                      callbacks::logger&) {
    for (size_t i = 0; i < z.p.size(); ++i) z.p[i] -= epsilon * z.g[i];
  }
  void update_q(point_t& z, Hamiltonian& h, double epsilon,
                callbacks::logger& logger) {
    h.update_q(z, epsilon);
    h.update_potential_gradient(z, logger);
  }
  void end_update_p(point_t& z, Hamiltonian& h, double epsilon,
                    callbacks::logger&) {
    for (size_t i = 0; i < z.p.size(); ++i) z.p[i] -= epsilon * z.g[i];
  }
  void evolve(point_t& z, Hamiltonian& h, double epsilon,
              callbacks::logger& logger) {
    begin_update_p(z, h, 0.5 * epsilon, logger);
    update_q(z, h, epsilon, logger);
    end_update_p(z, h, 0.5 * epsilon, logger);
  }
};

}  // namespace mcmc
}  // namespace stanx
#endif
