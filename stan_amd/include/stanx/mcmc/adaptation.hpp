// Warm-up adaptation of the host samplers: the step size by dual averaging
// and the inverse metric by running moments over the windows of a
// window_plan.  The arithmetic follows the reference so that draws match
// (paths relative to the stan-dev/stan 2.35.0 tree):
//   dual averaging       src/stan/mcmc/stepsize_adaptation.hpp:55-73
//   variance window      src/stan/mcmc/var_adaptation.hpp:17-46
//   covariance window    src/stan/mcmc/covar_adaptation.hpp:17-47
//   window schedule      src/stan/mcmc/windowed_adaptation.hpp:23-110 (warmup_plan.hpp)
// The running moments are Welford's (1962) update, the algorithm of Stan
// Math's welford_var_estimator and welford_covar_estimator (absent from the
// tree).  The device sampler applies the same updates per chain
// (csrc/nuts_batch.hip) and reads the same window_plan.
#ifndef STANX_MCMC_ADAPTATION_HPP
#define STANX_MCMC_ADAPTATION_HPP

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../callbacks.hpp"
#include "warmup_plan.hpp"

namespace stanx {
namespace mcmc {

// Nesterov dual averaging of log(step size) (Hoffman and Gelman 2014,
// algorithm 5).  Targets mean acceptance `delta`; `mu` is the point the
// iterates shrink towards, `gamma` the shrinkage, `t0` damps the first
// iterations and `kappa` sets the decay of the averaging weights.
class step_size_tuner {
 public:
  struct params {
    double mu = 0.5, delta = 0.5, gamma = 0.05, kappa = 0.75, t0 = 10;
  };

  // Out-of-domain values keep the current setting (delta in (0, 1); gamma,
  // kappa, t0 positive), as the reference's setters do.
  void set_mu(double m) { p_.mu = m; }
  void set_delta(double d) { p_.delta = (d > 0 && d < 1) ? d : p_.delta; }
  void set_gamma(double g) { p_.gamma = g > 0 ? g : p_.gamma; }
  void set_kappa(double k) { p_.kappa = k > 0 ? k : p_.kappa; }
  void set_t0(double t) { p_.t0 = t > 0 ? t : p_.t0; }
  const params& settings() const { return p_; }

  // forget the iterates (a new metric starts a new averaging run)
  void restart() { n_ = gap_ = avg_ = 0.0; }

  // One iteration with acceptance statistic `accept`: the step size to use
  // next.  gap_ is the running mean of (delta - accept), avg_ the weighted
  // mean of the log step sizes.
  double update(double accept) {
    n_ += 1.0;
    const double a = std::min(accept, 1.0);
    const double w = 1.0 / (n_ + p_.t0);
    gap_ = (1.0 - w) * gap_ + w * (p_.delta - a);
    const double log_eps = p_.mu - gap_ * std::sqrt(n_) / p_.gamma;
    const double v = std::pow(n_, -p_.kappa);
    avg_ = (1.0 - v) * avg_ + v * log_eps;
    return std::exp(log_eps);
  }
  // the step size after warm-up: the average iterate
  double final_step_size() const { return std::exp(avg_); }

  // iteration state (tests set it to restate the reference's goldens)
  void set_iterates(double n, double gap, double avg) {
    n_ = n;
    gap_ = gap;
    avg_ = avg;
  }
  double iterations() const { return n_; }
  double gap() const { return gap_; }
  double average() const { return avg_; }

 private:
  params p_;
  double n_ = 0.0, gap_ = 0.0, avg_ = 0.0;
};

// Running mean and variance per coordinate (Welford).
class running_variance {
 public:
  explicit running_variance(int n) : mean_(n, 0.0), ss_(n, 0.0) {}
  void clear() {
    count_ = 0.0;
    std::fill(mean_.begin(), mean_.end(), 0.0);
    std::fill(ss_.begin(), ss_.end(), 0.0);
  }
  void push(const std::vector<double>& x) {
    count_ += 1.0;
    for (size_t i = 0; i < mean_.size(); ++i) {
      const double d = x[i] - mean_[i];
      mean_[i] += d / count_;
      ss_[i] += (x[i] - mean_[i]) * d;
    }
  }
  double count() const { return count_; }
  // sample variance into `out` (left as it is with fewer than two draws)
  void estimate(std::vector<double>& out) const {
    if (count_ <= 1) return;
    for (size_t i = 0; i < mean_.size(); ++i) out[i] = ss_[i] / (count_ - 1.0);
  }
  // shrink `est` towards 1e-3 (times the identity) by the draw count:
  // est <- n / (n + 5) est + 1e-3 * 5 / (n + 5) on the diagonal
  static bool regularise(std::vector<double>& est, double n, int /*dim*/) {
    const double keep = n / (n + 5.0), floor = 1e-3 * (5.0 / (n + 5.0));
    bool ok = true;
    for (double& v : est) {
      v = keep * v + floor * 1.0;
      ok = ok && std::isfinite(v);
    }
    return ok;
  }

 private:
  double count_ = 0.0;
  std::vector<double> mean_, ss_;
};

// Running mean and covariance (Welford), row-major dim x dim:
// d = x - mean; mean += d / n; ss += (x - mean) d^T.
class running_covariance {
 public:
  explicit running_covariance(int n)
      : dim_(n), mean_(n, 0.0), ss_(static_cast<size_t>(n) * n, 0.0), d_(n, 0.0) {}
  void clear() {
    count_ = 0.0;
    std::fill(mean_.begin(), mean_.end(), 0.0);
    std::fill(ss_.begin(), ss_.end(), 0.0);
  }
  void push(const std::vector<double>& x) {
    count_ += 1.0;
    for (int i = 0; i < dim_; ++i) d_[i] = x[i] - mean_[i];
    for (int i = 0; i < dim_; ++i) mean_[i] += d_[i] / count_;
    for (int i = 0; i < dim_; ++i) {
      const double e = x[i] - mean_[i];
      double* row = ss_.data() + static_cast<size_t>(i) * dim_;
      for (int j = 0; j < dim_; ++j) row[j] += e * d_[j];
    }
  }
  double count() const { return count_; }
  void estimate(std::vector<double>& out) const {
    if (count_ <= 1) return;
    for (size_t k = 0; k < ss_.size(); ++k) out[k] = ss_[k] / (count_ - 1.0);
  }
  static bool regularise(std::vector<double>& est, double n, int dim) {
    const double keep = n / (n + 5.0), floor = 1e-3 * (5.0 / (n + 5.0));
    bool ok = true;
    for (int i = 0; i < dim; ++i)
      for (int j = 0; j < dim; ++j) {
        double& v = est[static_cast<size_t>(i) * dim + j];
        v = keep * v + floor * (i == j ? 1.0 : 0.0);
        ok = ok && std::isfinite(v);
      }
    return ok;
  }

 private:
  int dim_;
  double count_ = 0.0;
  std::vector<double> mean_, ss_, d_;
};

// The inverse metric estimated over the windows of a plan.  observe() is
// called once per warm-up transition with its draw; it returns true when a
// window closed and `metric` holds the new estimate.
template <class Moments>
class metric_windows {
 public:
  metric_windows(int dim, std::string estimate)
      : dim_(dim), what_(std::move(estimate)), moments_(dim) {
    plan_ = window_plan{};
  }

  void set_window_params(unsigned num_warmup, unsigned init_buffer, unsigned term_buffer,
                         unsigned base_window, callbacks::logger& logger) {
    plan_ = plan_windows(num_warmup, init_buffer, term_buffer, base_window, what_, logger);
    it_ = 0;
    cursor_ = 0;
  }
  const window_plan& plan() const { return plan_; }

  bool observe(std::vector<double>& metric, const std::vector<double>& q) {
    const unsigned it = it_++;
    if (plan_feeds(plan_, it)) moments_.push(q);
    if (!plan_closes(plan_, cursor_, it)) return false;
    ++cursor_;
    moments_.estimate(metric);
    const double n = moments_.count();
    if (!Moments::regularise(metric, n, dim_))
      throw std::runtime_error(
          "Numerical overflow in metric adaptation: the " + what_ +
          " estimate is not finite.  The sampler met extreme values on the "
          "unconstrained space; the posterior may be improper or too wide.");
    moments_.clear();
    return true;
  }

 private:
  int dim_;
  std::string what_;
  Moments moments_;
  window_plan plan_;
  unsigned it_ = 0;
  int cursor_ = 0;
};

using var_adaptation = metric_windows<running_variance>;
using covar_adaptation = metric_windows<running_covariance>;

}  // namespace mcmc
}  // namespace stanx
#endif
