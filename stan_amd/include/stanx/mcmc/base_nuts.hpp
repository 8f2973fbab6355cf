// Multinomial NUTS with a Euclidean metric (diagonal or dense), written as an
// explicit-stack (iterative) tree builder so that the same state machine
// can run one leapfrog step at a time; the GPU batched sampler uses the
// same frame layout per chain.  Semantics follow the recursive reference
// exactly, including the draw order of the RNG:
//   base_hmc                  src/stan/mcmc/hmc/base_hmc.hpp:28-215
//   base_nuts::transition     src/stan/mcmc/hmc/nuts/base_nuts.hpp:78-204
//   base_nuts::build_tree     src/stan/mcmc/hmc/nuts/base_nuts.hpp:247-352
//   compute_criterion         src/stan/mcmc/hmc/nuts/base_nuts.hpp:222-226
// Per merge of two half-subtrees: a log-sum-exp of their weights, a
// uniform progressive sample (one uniform draw whenever the final half
// does not dominate), and three U-turn checks (merged, left join, right
// join).  At the top level the sample is biased-progressive.
#ifndef STANX_MCMC_BASE_NUTS_HPP
#define STANX_MCMC_BASE_NUTS_HPP

#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../callbacks.hpp"
#include "../philox.hpp"
#include "hamiltonian.hpp"

namespace stanx {
namespace mcmc {

// log(1 + exp(a)) and log(exp(a) + exp(b)) with the -inf short cut of
// Stan Math's log_sum_exp(a, b).
inline double log1p_exp(double a) {
  return a > 0.0 ? a + std::log1p(std::exp(-a)) : std::log1p(std::exp(a));
}
inline double log_sum_exp(double a, double b) {
  const double inf = std::numeric_limits<double>::infinity();
  if (a == -inf) return b;
  if (a == inf && b == inf) return inf;
  if (a > b) return a + log1p_exp(b - a);
  return b + log1p_exp(a - b);
}
inline double dot(const vec_t& a, const vec_t& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

// draw returned by a transition (src/stan/mcmc/sample.hpp:12-56)
class sample {
 public:
  sample(const vec_t& q, double log_prob, double stat)
      : cont_params_(q), log_prob_(log_prob), accept_stat_(stat) {}
  const vec_t& cont_params() const { return cont_params_; }
  double log_prob() const { return log_prob_; }
  double accept_stat() const { return accept_stat_; }
  static void get_sample_param_names(std::vector<std::string>& names) {
    names.push_back("lp__");
    names.push_back("accept_stat__");
  }
  void get_sample_params(std::vector<double>& values) const {
    values.push_back(log_prob_);
    values.push_back(accept_stat_);
  }

 private:
  vec_t cont_params_;
  double log_prob_, accept_stat_;
};

template <class Model, class Metric = diag_e_metric<Model>>
class base_nuts {
 public:
  using hamiltonian_t = Metric;
  using point_t = typename Metric::point_t;

  base_nuts(const Model& model, rng_t& rng)
      : z_(static_cast<int>(model.num_params_r())),
        hamiltonian_(model),
        rng_(rng),
        nom_epsilon_(0.1),
        epsilon_(0.1),
        epsilon_jitter_(0.0),
        depth_(0),
        max_depth_(5),
        max_deltaH_(1000),
        n_leapfrog_(0),
        divergent_(false),
        energy_(0) {}
  virtual ~base_nuts() {}

  // ---- base_hmc surface ----
  point_t& z() { return z_; }
  const point_t& z() const { return z_; }
  hamiltonian_t& hamiltonian() { return hamiltonian_; }
  void set_metric(const vec_t& inv_e_metric) { z_.set_metric(inv_e_metric); }
  void set_nominal_stepsize(double e) {
    if (e > 0) nom_epsilon_ = e;
  }
  double get_nominal_stepsize() const { return nom_epsilon_; }
  double get_current_stepsize() const { return epsilon_; }
  void set_stepsize_jitter(double j) {
    if (j > 0 && j < 1) epsilon_jitter_ = j;
  }
  double get_stepsize_jitter() const { return epsilon_jitter_; }
  void set_max_depth(int d) {
    if (d > 0) max_depth_ = d;
  }
  void set_max_delta(double d) { max_deltaH_ = d; }
  int get_max_depth() const { return max_depth_; }
  double get_max_delta() const { return max_deltaH_; }
  int depth() const { return depth_; }
  int n_leapfrog() const { return n_leapfrog_; }
  bool divergent() const { return divergent_; }
  double energy() const { return energy_; }

  void seed(const vec_t& q) { z_.q = q; }
  void init_hamiltonian(callbacks::logger& logger) { hamiltonian_.init(z_, logger); }

  void sample_stepsize() {
    epsilon_ = nom_epsilon_;
    if (epsilon_jitter_)
      epsilon_ *= 1.0 + epsilon_jitter_ * (2.0 * rng_.uniform() - 1.0);
  }

  // Step-size heuristic run before sampling and after each metric update
  // (the behaviour of base_hmc::init_stepsize, base_hmc.hpp:93-161).  A
  // probe refreshes the momentum at the starting point and takes one
  // leapfrog step of the nominal size; the first probe fixes the direction
  // (grow while the energy change stays above log 0.8, otherwise shrink),
  // and the size doubles or halves until a probe's energy change crosses
  // log 0.8.  Every probe starts from the same point and draws a fresh
  // momentum, so the random stream is consumed as the reference consumes it.
  void init_stepsize(callbacks::logger& logger) {
    if (z_.q.empty()) {  // no step size in a zero-dimensional space
      nom_epsilon_ = std::numeric_limits<double>::quiet_NaN();
      return;
    }
    if (nom_epsilon_ == 0 || nom_epsilon_ > 1e7 || std::isnan(nom_epsilon_)) return;
    const double log_accept = std::log(0.8);
    ps_point start(static_cast<int>(z_.q.size()));
    start.assign(z_);
    auto probe = [&]() {  // energy change of one step from `start`
      z_.assign(start);
      hamiltonian_.sample_p(z_, rng_);
      hamiltonian_.init(z_, logger);
      const double before = hamiltonian_.H(z_);
      integrator_.evolve(z_, hamiltonian_, nom_epsilon_, logger);
      const double after = hamiltonian_.H(z_);
      return before - (std::isnan(after) ? std::numeric_limits<double>::infinity() : after);
    };
    const bool grow = probe() > log_accept;
    for (;;) {
      const double dH = probe();
      const bool crossed = grow ? !(dH > log_accept) : !(dH < log_accept);
      if (crossed) break;
      nom_epsilon_ *= grow ? 2.0 : 0.5;
      if (nom_epsilon_ > 1e7)
        throw std::runtime_error(
            "Step-size search: the step size grew past 1e7 with the energy change still "
            "small; the posterior is improper.  Check the model.");
      if (nom_epsilon_ == 0)
        throw std::runtime_error(
            "Step-size search: the step size shrank to zero without an acceptable energy "
            "change; the log density may not be continuous.");
    }
    z_.assign(start);
  }

  // One multinomial NUTS transition (the behaviour of base_nuts::transition,
  // base_nuts.hpp:78-204).  The trajectory is kept as its two sides,
  // side 0 backward and side 1 forward, so that growing either way is one
  // code path: a doubling picks a side by a uniform draw, hands the whole
  // trajectory so far to the other side as its subtree, and builds a new
  // subtree of the current depth from the chosen side's outer point.  The
  // new subtree's proposal replaces the sample with probability
  // min(1, w_new / w_old) (biased progressive sampling); three U-turn
  // checks over the merged trajectory and its two joints decide whether
  // to double again.
  sample transition(const sample& init_sample, callbacks::logger& logger) {
    const int D = static_cast<int>(z_.q.size());
    sample_stepsize();
    seed(init_sample.cont_params());
    hamiltonian_.sample_p(z_, rng_);
    hamiltonian_.init(z_, logger);
    ensure_frames(D);
    trajectory& t = traj_;
    vec_t p_sharp(z_.p.size());
    hamiltonian_.dtau_dp(z_, p_sharp);
    for (side& e : t.sides) {
      e.z.assign(z_);
      e.p_out = z_.p;
      e.p_in = z_.p;
      e.ps_out = p_sharp;
      e.ps_in = p_sharp;
    }
    t.sample.assign(z_);
    t.propose.assign(z_);
    t.rho = z_.p;

    const double H0 = hamiltonian_.H(z_);
    double log_w = 0;  // log of the trajectory's total weight
    int n_leapfrog = 0;
    double sum_metro = 0;
    depth_ = 0;
    divergent_ = false;
    while (depth_ < max_depth_) {
      const int s = rng_.uniform() > 0.5 ? 1 : 0;
      side& grow = t.sides[s];
      side& other = t.sides[1 - s];
      // the trajectory so far becomes the other side's subtree
      other.rho = t.rho;
      other.p_in = grow.p_out;
      other.ps_in = grow.ps_out;
      std::fill(grow.rho.begin(), grow.rho.end(), 0.0);
      z_.assign(grow.z);
      double log_w_new = -std::numeric_limits<double>::infinity();
      const bool valid =
          build_tree(depth_, t.propose, grow.ps_in, grow.ps_out, grow.rho, grow.p_in,
                     grow.p_out, H0, s ? 1.0 : -1.0, n_leapfrog, log_w_new, sum_metro, logger);
      grow.z.assign(z_);
      if (!valid) break;
      ++depth_;
      if (log_w_new > log_w || rng_.uniform() < std::exp(log_w_new - log_w))
        t.sample.assign(t.propose);
      log_w = log_sum_exp(log_w, log_w_new);
      if (!t.no_u_turn()) break;
    }
    n_leapfrog_ = n_leapfrog;
    z_.assign(t.sample);
    energy_ = hamiltonian_.H(z_);
    return sample(z_.q, -z_.V, sum_metro / static_cast<double>(n_leapfrog));
  }

  // sampler parameters of a draw (base_nuts.hpp:206-220)
  void get_sampler_param_names(std::vector<std::string>& names) const {
    names.push_back("stepsize__");
    names.push_back("treedepth__");
    names.push_back("n_leapfrog__");
    names.push_back("divergent__");
    names.push_back("energy__");
  }
  void get_sampler_params(std::vector<double>& values) const {
    values.push_back(this->epsilon_);
    values.push_back(this->depth_);
    values.push_back(this->n_leapfrog_);
    values.push_back(this->divergent_);
    values.push_back(this->energy_);
  }

  // diagnostic columns: position, momentum and potential gradient of the
  // current point, named after the model's unconstrained parameters with
  // p_ and g_ prefixes (ps_point.hpp:25-44)
  void get_sampler_diagnostic_names(const std::vector<std::string>& model_names,
                                    std::vector<std::string>& names) const {
    for (const auto& n : model_names) names.push_back(n);
    for (const auto& n : model_names) names.push_back("p_" + n);
    for (const auto& n : model_names) names.push_back("g_" + n);
  }
  void get_sampler_diagnostics(std::vector<double>& values) const {
    values.insert(values.end(), z_.q.begin(), z_.q.end());
    values.insert(values.end(), z_.p.begin(), z_.p.end());
    values.insert(values.end(), z_.g.begin(), z_.g.end());
  }
  // the adapted tuning parameters as one record (base_hmc.hpp:70-76)
  void write_sampler_state_struct(callbacks::structured_writer& w) const {
    w.begin_record();
    w.write("stepsize", nom_epsilon_);
    w.write("metric_type", std::string(point_t::metric_type()));
    z_.write_inv_metric(w);
    w.end_record();
  }

  static bool compute_criterion(const vec_t& p_sharp_minus,
                                const vec_t& p_sharp_plus, const vec_t& rho) {
    return dot(p_sharp_plus, rho) > 0 && dot(p_sharp_minus, rho) > 0;
  }

  // build_tree with the reference's argument semantics, evaluated with an
  // explicit frame stack instead of recursion.
  bool build_tree(int depth, ps_point& z_propose, vec_t& p_sharp_beg,
                  vec_t& p_sharp_end, vec_t& rho, vec_t& p_beg, vec_t& p_end,
                  double H0, double sign, int& n_leapfrog,
                  double& log_sum_weight, double& sum_metro_prob,
                  callbacks::logger& logger) {
    const int D = static_cast<int>(z_.q.size());
    ensure_frames(D);
    if (static_cast<int>(frames_.size()) < depth + 1)
      frames_.resize(depth + 1, frame(D));
    // dest[l]: where the subtree of depth l being built writes its results
    std::vector<dest>& dst = dests_;
    dst.assign(depth + 1, dest());
    dst[depth] = dest{&z_propose, &p_sharp_beg, &p_sharp_end, &rho,
                      &p_beg,     &p_end,       &log_sum_weight};
    for (int l = depth; l >= 1; --l) start_frame(l, dst);

    while (true) {
      // leaf: one leapfrog step writes a depth-0 subtree into dst[0]
      bool valid = leaf(dst[0], H0, sign, n_leapfrog, sum_metro_prob, logger);
      int l = 1;
      bool descended = false;
      while (l <= depth) {
        frame& f = frames_[l];
        if (!valid) return false;
        if (f.stage == 0) {
          // initial half done: start the final half from the current point
          f.stage = 1;
          f.z_propose_final.assign(z_);
          f.log_sum_weight_final = -std::numeric_limits<double>::infinity();
          std::fill(f.rho_final.begin(), f.rho_final.end(), 0.0);
          dst[l - 1] = dest{&f.z_propose_final, &f.p_sharp_final_beg,
                            dst[l].p_sharp_end, &f.rho_final,
                            &f.p_final_beg,     dst[l].p_end,
                            &f.log_sum_weight_final};
          for (int m = l - 1; m >= 1; --m) start_frame(m, dst);
          descended = true;
          break;
        }
        // final half done: merge the two halves of level l
        valid = merge(f, dst[l]);
        ++l;
      }
      if (!descended) return valid;
    }
  }

 protected:
  struct dest {
    ps_point* z_propose = nullptr;
    vec_t* p_sharp_beg = nullptr;
    vec_t* p_sharp_end = nullptr;
    vec_t* rho = nullptr;
    vec_t* p_beg = nullptr;
    vec_t* p_end = nullptr;
    double* log_sum_weight = nullptr;
  };
  struct frame {
    explicit frame(int D)
        : stage(0),
          p_init_end(D),
          p_sharp_init_end(D),
          rho_init(D),
          z_propose_final(D),
          p_final_beg(D),
          p_sharp_final_beg(D),
          rho_final(D),
          rho_subtree(D) {}
    int stage;
    double log_sum_weight_init = 0, log_sum_weight_final = 0;
    vec_t p_init_end, p_sharp_init_end, rho_init;
    ps_point z_propose_final;
    vec_t p_final_beg, p_sharp_final_beg, rho_final, rho_subtree;
  };
  // One end of the trajectory: the outermost point, the momentum and its
  // metric image (p-sharp) there and at the inner end of the subtree last
  // added on this side, and that subtree's momentum sum.
  struct side {
    explicit side(int D) : z(D), p_out(D), ps_out(D), p_in(D), ps_in(D), rho(D) {}
    ps_point z;
    vec_t p_out, ps_out, p_in, ps_in, rho;
  };
  struct trajectory {
    explicit trajectory(int D) : sides{side(D), side(D)}, sample(D), propose(D), rho(D), tmp(D) {}
    side sides[2];         // 0: backward, 1: forward
    ps_point sample, propose;
    vec_t rho, tmp;        // momentum sum of the whole trajectory; scratch
    // No U-turn across the whole trajectory (outer ends) nor across either
    // joint (each side's subtree extended by the other side's inner point).
    bool no_u_turn() {
      const side& b = sides[0];
      const side& f = sides[1];
      for (size_t i = 0; i < rho.size(); ++i) rho[i] = b.rho[i] + f.rho[i];
      bool ok = compute_criterion(b.ps_out, f.ps_out, rho);
      for (size_t i = 0; i < rho.size(); ++i) tmp[i] = b.rho[i] + f.p_in[i];
      ok &= compute_criterion(b.ps_out, f.ps_in, tmp);
      for (size_t i = 0; i < rho.size(); ++i) tmp[i] = f.rho[i] + b.p_in[i];
      ok &= compute_criterion(b.ps_in, f.ps_out, tmp);
      return ok;
    }
  };

  void ensure_frames(int D) {
    if (traj_.rho.size() != static_cast<size_t>(D)) traj_ = trajectory(D);
    const size_t want = static_cast<size_t>(max_depth_ + 1);
    if (frames_.empty() || frames_[0].rho_init.size() != static_cast<size_t>(D))
      frames_.assign(want, frame(D));
    else if (frames_.size() < want)
      frames_.resize(want, frame(D));
  }

  // begin building the initial half of level l (depth l-1 subtree)
  void start_frame(int l, std::vector<dest>& dst) {
    frame& f = frames_[l];
    f.stage = 0;
    f.log_sum_weight_init = -std::numeric_limits<double>::infinity();
    std::fill(f.rho_init.begin(), f.rho_init.end(), 0.0);
    dst[l - 1] = dest{dst[l].z_propose, dst[l].p_sharp_beg, &f.p_sharp_init_end,
                      &f.rho_init,      dst[l].p_beg,       &f.p_init_end,
                      &f.log_sum_weight_init};
  }

  // base case of build_tree (base_nuts.hpp:253-282)
  bool leaf(const dest& d, double H0, double sign, int& n_leapfrog,
            double& sum_metro_prob, callbacks::logger& logger) {
    integrator_.evolve(z_, hamiltonian_, sign * epsilon_, logger);
    ++n_leapfrog;
    double h = hamiltonian_.H(z_);
    if (std::isnan(h)) h = std::numeric_limits<double>::infinity();
    if ((h - H0) > max_deltaH_) divergent_ = true;
    *d.log_sum_weight = log_sum_exp(*d.log_sum_weight, H0 - h);
    if (H0 - h > 0)
      sum_metro_prob += 1;
    else
      sum_metro_prob += std::exp(H0 - h);
    d.z_propose->assign(z_);
    hamiltonian_.dtau_dp(z_, *d.p_sharp_beg);
    *d.p_sharp_end = *d.p_sharp_beg;
    vec_t& rho = *d.rho;
    for (size_t i = 0; i < rho.size(); ++i) rho[i] += z_.p[i];
    *d.p_beg = z_.p;
    *d.p_end = *d.p_beg;
    return !divergent_;
  }

  // merge of the two halves of a subtree (base_nuts.hpp:321-351)
  bool merge(frame& f, const dest& d) {
    const size_t D = f.rho_init.size();
    const double log_sum_weight_subtree =
        log_sum_exp(f.log_sum_weight_init, f.log_sum_weight_final);
    *d.log_sum_weight = log_sum_exp(*d.log_sum_weight, log_sum_weight_subtree);
    if (f.log_sum_weight_final > log_sum_weight_subtree) {
      d.z_propose->assign(f.z_propose_final);
    } else {
      const double accept_prob =
          std::exp(f.log_sum_weight_final - log_sum_weight_subtree);
      if (rng_.uniform() < accept_prob) d.z_propose->assign(f.z_propose_final);
    }
    vec_t& rs = f.rho_subtree;
    vec_t& rho = *d.rho;
    for (size_t i = 0; i < D; ++i) rs[i] = f.rho_init[i] + f.rho_final[i];
    for (size_t i = 0; i < D; ++i) rho[i] += rs[i];
    bool persist = compute_criterion(*d.p_sharp_beg, *d.p_sharp_end, rs);
    for (size_t i = 0; i < D; ++i) rs[i] = f.rho_init[i] + f.p_final_beg[i];
    persist &= compute_criterion(*d.p_sharp_beg, f.p_sharp_final_beg, rs);
    for (size_t i = 0; i < D; ++i) rs[i] = f.rho_final[i] + f.p_init_end[i];
    persist &= compute_criterion(f.p_sharp_init_end, *d.p_sharp_end, rs);
    return persist;
  }

  point_t z_;
  expl_leapfrog<hamiltonian_t> integrator_;
  hamiltonian_t hamiltonian_;
  rng_t& rng_;
  double nom_epsilon_, epsilon_, epsilon_jitter_;
  int depth_, max_depth_;
  double max_deltaH_;
  int n_leapfrog_;
  bool divergent_;
  double energy_;
  std::vector<frame> frames_;
  std::vector<dest> dests_;
  trajectory traj_{0};
};

}  // namespace mcmc
}  // namespace stanx
#endif
