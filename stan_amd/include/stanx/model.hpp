// The model boundary: a model_base-shaped interface and the gradient
// helpers the sampler calls.  Mirrors, in names, argument meaning and error
// behaviour:
//   prob_grad            src/stan/model/prob_grad.hpp:19-84
//   model_base           src/stan/model/model_base.hpp:36-703 (the double
//                        std::vector overloads :404-544 and the metadata
//                        methods :57-180, :348-388)
//   log_prob_grad        src/stan/model/log_prob_grad.hpp:28-50
//   log_prob_propto      src/stan/model/log_prob_propto.hpp:32-52
//   gradient             src/stan/model/gradient.hpp:14-35
//   model_base_crtp      src/stan/model/model_base_crtp.hpp:73-301
// A model supplies its value and gradient through log_prob_grad_impl (the
// GPU-backed model does so with one fused kernel instead of an AD tape).
// The reverse-mode overloads of the reference take this library's own var
// (stanx/var.hpp): model_base answers them with one precomputed-gradient
// entry on the caller's tape, and model_base_crtp runs a model written in
// the generated style (a log_prob template over double and var) through
// that tape.
#ifndef STANX_MODEL_HPP
#define STANX_MODEL_HPP

#include <cmath>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "callbacks.hpp"
#include "io.hpp"
#include "philox.hpp"
#include "var.hpp"

namespace stanx {
namespace model {

class prob_grad {
 public:
  explicit prob_grad(size_t num_params_r) : num_params_r__(num_params_r) {}
  virtual ~prob_grad() {}
  inline size_t num_params_r() const { return num_params_r__; }
  inline size_t num_params_i() const { return 0; }

 protected:
  size_t num_params_r__;
};

class model_base : public prob_grad {
 public:
  explicit model_base(size_t num_params_r) : prob_grad(num_params_r) {}
  virtual ~model_base() {}

  virtual std::string model_name() const = 0;
  virtual std::vector<std::string> model_compile_info() const {
    return {"stanx gfx950"};
  }
  virtual void get_param_names(std::vector<std::string>& names,
                               bool include_tparams = true,
                               bool include_gqs = true) const = 0;
  virtual void get_dims(std::vector<std::vector<size_t>>& dimss,
                        bool include_tparams = true,
                        bool include_gqs = true) const = 0;
  virtual void constrained_param_names(std::vector<std::string>& names,
                                       bool include_tparams = true,
                                       bool include_gqs = true) const = 0;
  virtual void unconstrained_param_names(std::vector<std::string>& names,
                                         bool include_tparams = true,
                                         bool include_gqs = true) const = 0;

  // Value and gradient of log p(params_r) with or without constants and
  // the change-of-variables adjustment.  Throws std::domain_error when the
  // density cannot be evaluated (the reference's recoverable rejection).
  virtual double log_prob_grad_impl(const std::vector<double>& params_r,
                                    std::vector<double>* gradient,
                                    bool propto, bool jacobian,
                                    std::ostream* msgs) const = 0;

  // The four double overloads of the reference (:404-458).
  virtual double log_prob(std::vector<double>& params_r, std::vector<int>& params_i,
                          std::ostream* msgs) const {
    return log_prob_grad_impl(params_r, nullptr, false, false, msgs);
  }
  virtual double log_prob_jacobian(std::vector<double>& params_r,
                                   std::vector<int>& params_i, std::ostream* msgs) const {
    return log_prob_grad_impl(params_r, nullptr, false, true, msgs);
  }
  virtual double log_prob_propto(std::vector<double>& params_r,
                                 std::vector<int>& params_i, std::ostream* msgs) const {
    return log_prob_grad_impl(params_r, nullptr, true, false, msgs);
  }
  virtual double log_prob_propto_jacobian(std::vector<double>& params_r,
                                          std::vector<int>& params_i,
                                          std::ostream* msgs) const {
    return log_prob_grad_impl(params_r, nullptr, true, true, msgs);
  }
  // The four reverse-mode overloads (:460-544): the log density as one
  // tape entry whose edges to params_r carry the gradient.
  virtual math::var log_prob(std::vector<math::var>& params_r, std::vector<int>& params_i,
                             std::ostream* msgs) const {
    return var_log_prob(params_r, false, false, msgs);
  }
  virtual math::var log_prob_jacobian(std::vector<math::var>& params_r,
                                      std::vector<int>& params_i, std::ostream* msgs) const {
    return var_log_prob(params_r, false, true, msgs);
  }
  virtual math::var log_prob_propto(std::vector<math::var>& params_r,
                                    std::vector<int>& params_i, std::ostream* msgs) const {
    return var_log_prob(params_r, true, false, msgs);
  }
  virtual math::var log_prob_propto_jacobian(std::vector<math::var>& params_r,
                                             std::vector<int>& params_i,
                                             std::ostream* msgs) const {
    return var_log_prob(params_r, true, true, msgs);
  }
  // log_prob<propto, jacobian> dispatch to the overloads (model_base.hpp:306-340)
  template <bool propto, bool jacobian, typename T>
  T log_prob(std::vector<T>& params_r, std::vector<int>& params_i,
             std::ostream* msgs = nullptr) const {
    if (propto)
      return jacobian ? log_prob_propto_jacobian(params_r, params_i, msgs)
                      : log_prob_propto(params_r, params_i, msgs);
    return jacobian ? log_prob_jacobian(params_r, params_i, msgs)
                    : log_prob(params_r, params_i, msgs);
  }

  virtual void transform_inits(const io::var_context& context,
                               std::vector<int>& params_i,
                               std::vector<double>& params_r,
                               std::ostream* msgs) const = 0;
  virtual void write_array(rng_t& rng, std::vector<double>& params_r,
                           std::vector<int>& params_i,
                           std::vector<double>& vars,
                           bool include_tparams = true,
                           bool include_gqs = true,
                           std::ostream* msgs = nullptr) const = 0;
  virtual void unconstrain_array(const std::vector<double>& params_constrained,
                                 std::vector<double>& params_unconstrained,
                                 std::ostream* msgs = nullptr) const = 0;

 protected:
  math::var var_log_prob(std::vector<math::var>& params_r, bool propto, bool jacobian,
                         std::ostream* msgs) const {
    std::vector<double> x(params_r.size()), g(params_r.size());
    for (size_t i = 0; i < x.size(); ++i) x[i] = params_r[i].val();
    const double lp = log_prob_grad_impl(x, &g, propto, jacobian, msgs);
    return math::precomputed_gradients(lp, params_r, g);
  }
};

// A model written in the generated style: M defines
//   template <bool propto, bool jacobian, typename T>
//   T log_prob(std::vector<T>& params_r, std::vector<int>& params_i,
//              std::ostream* msgs) const;
// for T = double and T = math::var, plus the metadata methods.  Every
// virtual of model_base forwards to that template; the gradient comes from
// the var instantiation on this thread's tape (log_prob_grad.hpp:28-50),
// which is cleared afterwards, also when the model throws.
template <class M>
class model_base_crtp : public model_base {
 public:
  explicit model_base_crtp(size_t num_params_r) : model_base(num_params_r) {}

  double log_prob(std::vector<double>& theta, std::vector<int>& theta_i,
                  std::ostream* msgs) const override {
    return self().template log_prob<false, false>(theta, theta_i, msgs);
  }
  math::var log_prob(std::vector<math::var>& theta, std::vector<int>& theta_i,
                     std::ostream* msgs) const override {
    return self().template log_prob<false, false>(theta, theta_i, msgs);
  }
  double log_prob_jacobian(std::vector<double>& theta, std::vector<int>& theta_i,
                           std::ostream* msgs) const override {
    return self().template log_prob<false, true>(theta, theta_i, msgs);
  }
  math::var log_prob_jacobian(std::vector<math::var>& theta, std::vector<int>& theta_i,
                              std::ostream* msgs) const override {
    return self().template log_prob<false, true>(theta, theta_i, msgs);
  }
  double log_prob_propto(std::vector<double>& theta, std::vector<int>& theta_i,
                         std::ostream* msgs) const override {
    return self().template log_prob<true, false>(theta, theta_i, msgs);
  }
  math::var log_prob_propto(std::vector<math::var>& theta, std::vector<int>& theta_i,
                            std::ostream* msgs) const override {
    return self().template log_prob<true, false>(theta, theta_i, msgs);
  }
  double log_prob_propto_jacobian(std::vector<double>& theta, std::vector<int>& theta_i,
                                  std::ostream* msgs) const override {
    return self().template log_prob<true, true>(theta, theta_i, msgs);
  }
  math::var log_prob_propto_jacobian(std::vector<math::var>& theta,
                                     std::vector<int>& theta_i,
                                     std::ostream* msgs) const override {
    return self().template log_prob<true, true>(theta, theta_i, msgs);
  }

  double log_prob_grad_impl(const std::vector<double>& params_r, std::vector<double>* gradient,
                            bool propto, bool jacobian, std::ostream* msgs) const override {
    std::vector<int> params_i;
    if (!gradient) {
      std::vector<double> x(params_r);
      return propto ? (jacobian ? self().template log_prob<true, true>(x, params_i, msgs)
                                : self().template log_prob<true, false>(x, params_i, msgs))
                    : (jacobian ? self().template log_prob<false, true>(x, params_i, msgs)
                                : self().template log_prob<false, false>(x, params_i, msgs));
    }
    math::tape_scope scope;
    std::vector<math::var> v(params_r.begin(), params_r.end());
    const math::var lp =
        propto ? (jacobian ? self().template log_prob<true, true>(v, params_i, msgs)
                           : self().template log_prob<true, false>(v, params_i, msgs))
               : (jacobian ? self().template log_prob<false, true>(v, params_i, msgs)
                           : self().template log_prob<false, false>(v, params_i, msgs));
    lp.grad();
    gradient->resize(v.size());
    for (size_t i = 0; i < v.size(); ++i) (*gradient)[i] = v[i].adj();
    return lp.val();
  }

 private:
  const M& self() const { return *static_cast<const M*>(this); }
};

// stan::model::log_prob_grad<propto, jacobian> (log_prob_grad.hpp:28-50):
// returns the log density, writes the gradient (resized by the callee).
template <bool propto, bool jacobian_adjust_transform, class M>
double log_prob_grad(const M& model, std::vector<double>& params_r,
                     std::vector<int>& params_i,
                     std::vector<double>& gradient,
                     std::ostream* msgs = nullptr) {
  gradient.resize(model.num_params_r());
  return model.log_prob_grad_impl(params_r, &gradient, propto,
                                  jacobian_adjust_transform, msgs);
}

// stan::model::log_prob_propto<jacobian> (log_prob_propto.hpp:32-52)
template <bool jacobian_adjust_transform, class M>
double log_prob_propto(const M& model, std::vector<double>& params_r,
                       std::vector<int>& params_i,
                       std::ostream* msgs = nullptr) {
  return model.log_prob_grad_impl(params_r, nullptr, true,
                                  jacobian_adjust_transform, msgs);
}

// stan::model::gradient (gradient.hpp:14-35): value f and gradient grad_f
// of log_prob<propto=true, jacobian=true>; model messages go to the logger
// and exceptions are rethrown.
template <class M>
void gradient(const M& model, const std::vector<double>& x, double& f,
              std::vector<double>& grad_f, callbacks::logger& logger) {
  std::stringstream ss;
  try {
    grad_f.resize(model.num_params_r());
    f = model.log_prob_grad_impl(x, &grad_f, true, true, &ss);
  } catch (std::exception& e) {
    if (ss.str().length() > 0) logger.info(ss);
    throw;
  }
  if (ss.str().length() > 0) logger.info(ss);
}

// A model given as a plain C callback (value + gradient of the propto,
// jacobian log density); parameters are named q.1 .. q.D.
typedef int (*lpgrad_fn)(void* ctx, const double* q, int D, double* lp,
                         double* grad);

class callback_model : public model_base {
 public:
  callback_model(lpgrad_fn f, void* ctx, int D)
      : model_base(static_cast<size_t>(D)), f_(f), ctx_(ctx) {}
  std::string model_name() const override { return "callback_model"; }
  void get_param_names(std::vector<std::string>& names, bool,
                       bool) const override {
    names.assign(1, "q");
  }
  void get_dims(std::vector<std::vector<size_t>>& dimss, bool,
                bool) const override {
    dimss.assign(1, std::vector<size_t>{num_params_r()});
  }
  void constrained_param_names(std::vector<std::string>& names, bool,
                               bool) const override {
    for (size_t i = 0; i < num_params_r(); ++i)
      names.push_back("q." + std::to_string(i + 1));
  }
  void unconstrained_param_names(std::vector<std::string>& names, bool,
                                 bool) const override {
    constrained_param_names(names, true, true);
  }
  double log_prob_grad_impl(const std::vector<double>& q,
                            std::vector<double>* gradient, bool, bool,
                            std::ostream*) const override {
    std::vector<double> g(num_params_r());
    double lp = 0.0;
    int st = f_(ctx_, q.data(), static_cast<int>(q.size()), &lp, g.data());
    if (st == 1) throw std::domain_error("model callback: domain error");
    if (st != 0) throw std::runtime_error("model callback: fatal error");
    if (gradient) *gradient = g;
    return lp;
  }
  void transform_inits(const io::var_context& context, std::vector<int>&,
                       std::vector<double>& params_r,
                       std::ostream*) const override {
    params_r = context.vals_r("q");
  }
  void write_array(rng_t&, std::vector<double>& params_r, std::vector<int>&,
                   std::vector<double>& vars, bool, bool,
                   std::ostream*) const override {
    vars = params_r;
  }
  void unconstrain_array(const std::vector<double>& c,
                         std::vector<double>& u,
                         std::ostream*) const override {
    u = c;
  }

 private:
  lpgrad_fn f_;
  void* ctx_;
};

}  // namespace model
}  // namespace stanx
#endif
