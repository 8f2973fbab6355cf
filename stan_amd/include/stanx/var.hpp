// A minimal reverse-mode scalar for the model boundary.
//
// The reference's model_base declares log_prob overloads over Stan Math's
// reverse-mode type (src/stan/model/model_base.hpp:191-302, 404-544) and
// log_prob_grad differentiates a model through it
// (src/stan/model/log_prob_grad.hpp:28-50).  Stan Math is not part of this
// library, so it carries its own: a Wengert list per thread.  Every var is
// an index into the calling thread's tape; a tape entry holds the value,
// the adjoint and the contiguous run of its edges (operand index, partial
// derivative).  grad() seeds one entry with adjoint 1 and sweeps the tape
// backwards once.
//
// The GPU models enter a tape as ONE entry whose edges carry the gradient
// the fused kernel computed (precomputed_gradients): a log density over a
// million rows costs one entry and D edges, not a node per row.
#ifndef STANX_VAR_HPP
#define STANX_VAR_HPP

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

namespace stanx {
namespace math {

struct tape {
  std::vector<double> val, adj;
  std::vector<uint32_t> first;  // entry i's edges: [first[i], first[i + 1])
  std::vector<uint32_t> to;     // operand entry of each edge
  std::vector<double> d;        // partial derivative of each edge
  tape() { first.push_back(0); }
  uint32_t push(double v) {
    val.push_back(v);
    adj.push_back(0.0);
    first.push_back(static_cast<uint32_t>(to.size()));
    return static_cast<uint32_t>(val.size() - 1);
  }
  void edge(uint32_t operand, double partial) {
    to.push_back(operand);
    d.push_back(partial);
    first.back() = static_cast<uint32_t>(to.size());
  }
  void clear() {
    val.clear();
    adj.clear();
    to.clear();
    d.clear();
    first.assign(1, 0);
  }
};

inline tape& active_tape() {
  thread_local tape t;
  return t;
}

class var {
 public:
  var() : id_(active_tape().push(0.0)) {}
  var(double v) : id_(active_tape().push(v)) {}  // NOLINT: constants convert
  var(int v) : id_(active_tape().push(static_cast<double>(v))) {}  // NOLINT
  static var entry(uint32_t id) {
    var x(tag{});
    x.id_ = id;
    return x;
  }
  double val() const { return active_tape().val[id_]; }
  double adj() const { return active_tape().adj[id_]; }
  uint32_t id() const { return id_; }

  // d(this)/d(entry) for every entry of the tape, into the adjoints
  void grad() const {
    tape& t = active_tape();
    std::fill(t.adj.begin(), t.adj.end(), 0.0);
    t.adj[id_] = 1.0;
    for (int64_t i = id_; i >= 0; --i) {
      const double a = t.adj[i];
      if (a == 0.0) continue;
      for (uint32_t e = t.first[i]; e < t.first[i + 1]; ++e) t.adj[t.to[e]] += t.d[e] * a;
    }
  }

  var& operator+=(const var& b);
  var& operator-=(const var& b);
  var& operator*=(const var& b);
  var& operator/=(const var& b);

 private:
  struct tag {};
  explicit var(tag) : id_(0) {}
  uint32_t id_;
};

// An entry of value v with edges to a (and b).
inline var unary(double v, const var& a, double da) {
  tape& t = active_tape();
  const uint32_t id = t.push(v);
  t.edge(a.id(), da);
  return var::entry(id);
}
inline var binary(double v, const var& a, double da, const var& b, double db) {
  tape& t = active_tape();
  const uint32_t id = t.push(v);
  t.edge(a.id(), da);
  t.edge(b.id(), db);
  return var::entry(id);
}

// One entry for a value whose gradient with respect to `operands` was
// computed elsewhere (the GPU log density).
inline var precomputed_gradients(double value, const std::vector<var>& operands,
                                 const std::vector<double>& gradients) {
  if (operands.size() != gradients.size())
    throw std::invalid_argument("precomputed_gradients: operand and gradient sizes differ");
  tape& t = active_tape();
  const uint32_t id = t.push(value);
  for (size_t i = 0; i < operands.size(); ++i) t.edge(operands[i].id(), gradients[i]);
  return var::entry(id);
}

// Forget every entry of this thread's tape (the reference's recover_memory).
inline void recover_memory() { active_tape().clear(); }

// Clears the tape when it goes out of scope, also on an exception.
struct tape_scope {
  tape_scope() = default;
  tape_scope(const tape_scope&) = delete;
  tape_scope& operator=(const tape_scope&) = delete;
  ~tape_scope() { recover_memory(); }
};

inline double value_of(const var& x) { return x.val(); }
inline double value_of(double x) { return x; }

inline var operator+(const var& a, const var& b) { return binary(a.val() + b.val(), a, 1.0, b, 1.0); }
inline var operator-(const var& a, const var& b) { return binary(a.val() - b.val(), a, 1.0, b, -1.0); }
inline var operator*(const var& a, const var& b) {
  return binary(a.val() * b.val(), a, b.val(), b, a.val());
}
inline var operator/(const var& a, const var& b) {
  const double q = a.val() / b.val();
  return binary(q, a, 1.0 / b.val(), b, -q / b.val());
}
inline var operator-(const var& a) { return unary(-a.val(), a, -1.0); }
inline var operator+(const var& a, double b) { return unary(a.val() + b, a, 1.0); }
inline var operator+(double a, const var& b) { return unary(a + b.val(), b, 1.0); }
inline var operator-(const var& a, double b) { return unary(a.val() - b, a, 1.0); }
inline var operator-(double a, const var& b) { return unary(a - b.val(), b, -1.0); }
inline var operator*(const var& a, double b) { return unary(a.val() * b, a, b); }
inline var operator*(double a, const var& b) { return unary(a * b.val(), b, a); }
inline var operator/(const var& a, double b) { return unary(a.val() / b, a, 1.0 / b); }
inline var operator/(double a, const var& b) {
  const double q = a / b.val();
  return unary(q, b, -q / b.val());
}
inline var& var::operator+=(const var& b) { return *this = *this + b; }
inline var& var::operator-=(const var& b) { return *this = *this - b; }
inline var& var::operator*=(const var& b) { return *this = *this * b; }
inline var& var::operator/=(const var& b) { return *this = *this / b; }

inline var exp(const var& a) {
  const double e = std::exp(a.val());
  return unary(e, a, e);
}
inline var log(const var& a) { return unary(std::log(a.val()), a, 1.0 / a.val()); }
inline var log1p(const var& a) { return unary(std::log1p(a.val()), a, 1.0 / (1.0 + a.val())); }
inline var sqrt(const var& a) {
  const double s = std::sqrt(a.val());
  return unary(s, a, 0.5 / s);
}
inline var square(const var& a) { return unary(a.val() * a.val(), a, 2.0 * a.val()); }
inline double square(double a) { return a * a; }
inline var pow(const var& a, double b) {
  return unary(std::pow(a.val(), b), a, b * std::pow(a.val(), b - 1.0));
}

inline bool operator<(const var& a, const var& b) { return a.val() < b.val(); }
inline bool operator>(const var& a, const var& b) { return a.val() > b.val(); }
inline bool operator<(const var& a, double b) { return a.val() < b; }
inline bool operator>(const var& a, double b) { return a.val() > b; }

}  // namespace math
}  // namespace stanx
#endif
