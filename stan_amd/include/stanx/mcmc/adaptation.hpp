// Warm-up adaptation, mirroring
//   stepsize_adaptation   src/stan/mcmc/stepsize_adaptation.hpp:11-84
//   windowed_adaptation   src/stan/mcmc/windowed_adaptation.hpp:12-123
//   var_adaptation        src/stan/mcmc/var_adaptation.hpp:12-50
//   covar_adaptation      src/stan/mcmc/covar_adaptation.hpp:12-51
// The running variance and covariance are Welford's (1962) update, the
// algorithm of Stan Math's welford_var_estimator and
// welford_covar_estimator (absent from the tree).
#ifndef STANX_MCMC_ADAPTATION_HPP
#define STANX_MCMC_ADAPTATION_HPP

#include <cmath>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../callbacks.hpp"

namespace stanx {
namespace mcmc {

class base_adaptation {
 public:
  virtual ~base_adaptation() {}
  virtual void restart() {}
};

class stepsize_adaptation : public base_adaptation {
 public:
  stepsize_adaptation()
      : mu_(0.5), delta_(0.5), gamma_(0.05), kappa_(0.75), t0_(10) {
    restart();
  }
  void set_mu(double m) { mu_ = m; }
  void set_delta(double d) {
    if (d > 0 && d < 1) delta_ = d;
  }
  void set_gamma(double g) {
    if (g > 0) gamma_ = g;
  }
  void set_kappa(double k) {
    if (k > 0) kappa_ = k;
  }
  void set_t0(double t) {
    if (t > 0) t0_ = t;
  }
  double get_mu() const { return mu_; }
  double get_delta() const { return delta_; }
  double get_gamma() const { return gamma_; }
  double get_kappa() const { return kappa_; }
  double get_t0() const { return t0_; }
  void restart() override {
    counter_ = 0;
    s_bar_ = 0;
    x_bar_ = 0;
  }
  // Nesterov dual averaging of log(epsilon) (Hoffman and Gelman 2014)
  void learn_stepsize(double& epsilon, double adapt_stat) {
    ++counter_;
    adapt_stat = adapt_stat > 1 ? 1 : adapt_stat;
    const double eta = 1.0 / (counter_ + t0_);
    s_bar_ = (1.0 - eta) * s_bar_ + eta * (delta_ - adapt_stat);
    const double x = mu_ - s_bar_ * std::sqrt(counter_) / gamma_;
    const double x_eta = std::pow(counter_, -kappa_);
    x_bar_ = (1.0 - x_eta) * x_bar_ + x_eta * x;
    epsilon = std::exp(x);
  }
  void complete_adaptation(double& epsilon) { epsilon = std::exp(x_bar_); }

  // full state access (tests restate goldens that set it directly)
  void set_state(double counter, double s_bar, double x_bar) {
    counter_ = counter;
    s_bar_ = s_bar;
    x_bar_ = x_bar;
  }
  double counter() const { return counter_; }
  double s_bar() const { return s_bar_; }
  double x_bar() const { return x_bar_; }

 protected:
  double counter_, s_bar_, x_bar_;
  double mu_, delta_, gamma_, kappa_, t0_;
};

class windowed_adaptation : public base_adaptation {
 public:
  explicit windowed_adaptation(std::string name) : estimator_name_(name) {
    num_warmup_ = 0;
    adapt_init_buffer_ = 0;
    adapt_term_buffer_ = 0;
    adapt_base_window_ = 0;
    restart();
  }
  void restart() override {
    adapt_window_counter_ = 0;
    adapt_window_size_ = adapt_base_window_;
    adapt_next_window_ = adapt_init_buffer_ + adapt_window_size_ - 1;
  }
  void set_window_params(unsigned int num_warmup, unsigned int init_buffer,
                         unsigned int term_buffer, unsigned int base_window,
                         callbacks::logger& logger) {
    if (num_warmup < 20) {
      logger.info("WARNING: No " + estimator_name_ + " estimation is");
      logger.info("         performed for num_warmup < 20");
      logger.info("");
      return;
    }
    if (init_buffer + base_window + term_buffer > num_warmup) {
      logger.info("WARNING: There aren't enough warmup iterations to fit the");
      logger.info("         three stages of adaptation as currently configured.");
      num_warmup_ = num_warmup;
      adapt_init_buffer_ = 0.15 * num_warmup;
      adapt_term_buffer_ = 0.10 * num_warmup;
      adapt_base_window_ = num_warmup - (adapt_init_buffer_ + adapt_term_buffer_);
      logger.info("         Reducing each adaptation stage to 15%/75%/10% of");
      logger.info("         the given number of warmup iterations:");
      std::stringstream m1, m2, m3;
      m1 << "           init_buffer = " << adapt_init_buffer_;
      logger.info(m1);
      m2 << "           adapt_window = " << adapt_base_window_;
      logger.info(m2);
      m3 << "           term_buffer = " << adapt_term_buffer_;
      logger.info(m3);
      logger.info("");
      return;
    }
    num_warmup_ = num_warmup;
    adapt_init_buffer_ = init_buffer;
    adapt_term_buffer_ = term_buffer;
    adapt_base_window_ = base_window;
    restart();
  }
  bool adaptation_window() const {
    return (adapt_window_counter_ >= adapt_init_buffer_) &&
           (adapt_window_counter_ < num_warmup_ - adapt_term_buffer_) &&
           (adapt_window_counter_ != num_warmup_);
  }
  bool end_adaptation_window() const {
    return (adapt_window_counter_ == adapt_next_window_) &&
           (adapt_window_counter_ != num_warmup_);
  }
  void compute_next_window() {
    if (adapt_next_window_ == num_warmup_ - adapt_term_buffer_ - 1) return;
    adapt_window_size_ *= 2;
    adapt_next_window_ = adapt_window_counter_ + adapt_window_size_;
    if (adapt_next_window_ == num_warmup_ - adapt_term_buffer_ - 1) return;
    unsigned int next_window_boundary =
        adapt_next_window_ + 2 * adapt_window_size_;
    if (next_window_boundary >= num_warmup_ - adapt_term_buffer_)
      adapt_next_window_ = num_warmup_ - adapt_term_buffer_ - 1;
  }
  unsigned int window_counter() const { return adapt_window_counter_; }
  unsigned int next_window() const { return adapt_next_window_; }
  unsigned int init_buffer() const { return adapt_init_buffer_; }
  unsigned int term_buffer() const { return adapt_term_buffer_; }
  unsigned int base_window() const { return adapt_base_window_; }

 protected:
  std::string estimator_name_;
  unsigned int num_warmup_, adapt_init_buffer_, adapt_term_buffer_,
      adapt_base_window_;
  unsigned int adapt_window_counter_, adapt_next_window_, adapt_window_size_;
};

class welford_var_estimator {
 public:
  explicit welford_var_estimator(int n) : m_(n, 0.0), m2_(n, 0.0) {}
  void restart() {
    num_samples_ = 0;
    std::fill(m_.begin(), m_.end(), 0.0);
    std::fill(m2_.begin(), m2_.end(), 0.0);
  }
  void add_sample(const std::vector<double>& q) {
    num_samples_ += 1.0;
    for (size_t i = 0; i < m_.size(); ++i) {
      const double delta = q[i] - m_[i];
      m_[i] += delta / num_samples_;
      m2_[i] += (q[i] - m_[i]) * delta;
    }
  }
  void sample_variance(std::vector<double>& var) const {
    if (num_samples_ > 1)
      for (size_t i = 0; i < m_.size(); ++i) var[i] = m2_[i] / (num_samples_ - 1.0);
  }
  double num_samples() const { return num_samples_; }

 private:
  double num_samples_ = 0;
  std::vector<double> m_, m2_;
};

class var_adaptation : public windowed_adaptation {
 public:
  explicit var_adaptation(int n)
      : windowed_adaptation("variance"), estimator_(n) {}
  bool learn_variance(std::vector<double>& var, const std::vector<double>& q) {
    if (adaptation_window()) estimator_.add_sample(q);
    if (end_adaptation_window()) {
      compute_next_window();
      estimator_.sample_variance(var);
      const double n = estimator_.num_samples();
      bool finite = true;
      for (auto& v : var) {
        v = (n / (n + 5.0)) * v + 1e-3 * (5.0 / (n + 5.0)) * 1.0;
        if (!std::isfinite(v)) finite = false;
      }
      if (!finite)
        throw std::runtime_error(
            "Numerical overflow in metric adaptation. "
            "This occurs when the sampler encounters extreme values on the "
            "unconstrained space; this may happen when the posterior density "
            "function is too wide or improper. "
            "There may be problems with your model specification.");
      estimator_.restart();
      ++adapt_window_counter_;
      return true;
    }
    ++adapt_window_counter_;
    return false;
  }

 protected:
  welford_var_estimator estimator_;
};

// Running mean and covariance: delta = q - m, m += delta / n,
// m2 += (q - m) delta^T (row-major D x D), covariance m2 / (n - 1).
class welford_covar_estimator {
 public:
  explicit welford_covar_estimator(int n)
      : n_(n), m_(n, 0.0), m2_(static_cast<size_t>(n) * n, 0.0), delta_(n, 0.0) {}
  void restart() {
    num_samples_ = 0;
    std::fill(m_.begin(), m_.end(), 0.0);
    std::fill(m2_.begin(), m2_.end(), 0.0);
  }
  void add_sample(const std::vector<double>& q) {
    num_samples_ += 1.0;
    for (int i = 0; i < n_; ++i) delta_[i] = q[i] - m_[i];
    for (int i = 0; i < n_; ++i) m_[i] += delta_[i] / num_samples_;
    for (int i = 0; i < n_; ++i) {
      const double d = q[i] - m_[i];
      double* row = m2_.data() + static_cast<size_t>(i) * n_;
      for (int j = 0; j < n_; ++j) row[j] += d * delta_[j];
    }
  }
  void sample_covariance(std::vector<double>& covar) const {
    if (num_samples_ > 1)
      for (size_t k = 0; k < m2_.size(); ++k) covar[k] = m2_[k] / (num_samples_ - 1.0);
  }
  double num_samples() const { return num_samples_; }

 private:
  int n_;
  double num_samples_ = 0;
  std::vector<double> m_, m2_, delta_;
};

// Windowed covariance estimation of a dense inverse metric (row-major):
// at each window end the regularised estimate n/(n+5) C + 1e-3 5/(n+5) I.
class covar_adaptation : public windowed_adaptation {
 public:
  explicit covar_adaptation(int n)
      : windowed_adaptation("covariance"), n_(n), estimator_(n) {}
  bool learn_covariance(std::vector<double>& covar, const std::vector<double>& q) {
    if (adaptation_window()) estimator_.add_sample(q);
    if (end_adaptation_window()) {
      compute_next_window();
      estimator_.sample_covariance(covar);
      const double n = estimator_.num_samples();
      const double a = n / (n + 5.0), b = 1e-3 * (5.0 / (n + 5.0));
      bool finite = true;
      for (int i = 0; i < n_; ++i)
        for (int j = 0; j < n_; ++j) {
          double& c = covar[static_cast<size_t>(i) * n_ + j];
          c = a * c + b * (i == j ? 1.0 : 0.0);
          if (!std::isfinite(c)) finite = false;
        }
      if (!finite)
        throw std::runtime_error(
            "Numerical overflow in metric adaptation. "
            "This occurs when the sampler encounters extreme values on the "
            "unconstrained space; this may happen when the posterior density "
            "function is too wide or improper. "
            "There may be problems with your model specification.");
      estimator_.restart();
      ++adapt_window_counter_;
      return true;
    }
    ++adapt_window_counter_;
    return false;
  }

 protected:
  int n_;
  welford_covar_estimator estimator_;
};

}  // namespace mcmc
}  // namespace stanx
#endif
