// Data files in Stan's JSON format as a var_context, after
// src/stan/io/json/json_data.hpp:42-73 (json_data, built by
// json_data_handler.hpp).  A file is one object whose members are numbers
// or rectangular nested arrays of numbers; "NaN", "Inf", "-Inf",
// "Infinity" and "-Infinity" strings are the non-finite reals.  Arrays are
// written row-major in JSON and stored column-major (first index fastest),
// the layout of var_context values, with their dimensions.
#ifndef STANX_JSON_DATA_HPP
#define STANX_JSON_DATA_HPP

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "io.hpp"

namespace stanx {
namespace io {

class json_data : public array_var_context {
 public:
  explicit json_data(const std::string& text) : s_(text) { parse(); }
  static json_data from_file(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return json_data(ss.str());
  }
  // true when every value of the variable is an integer literal
  bool is_int(const std::string& name) const {
    for (const auto& v : ints_)
      if (v == name) return true;
    return false;
  }
  const std::vector<std::string>& names() const { return names_; }

 private:
  struct node {
    bool leaf = true, integer = true;
    double value = 0;
    std::vector<node> items;
  };

  void fail(const std::string& what) const {
    throw std::runtime_error("JSON data: " + what + " at offset " + std::to_string(p_));
  }
  void ws() {
    while (p_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[p_]))) ++p_;
  }
  bool eat(char c) {
    ws();
    if (p_ < s_.size() && s_[p_] == c) {
      ++p_;
      return true;
    }
    return false;
  }
  std::string str() {
    ws();
    if (p_ >= s_.size() || s_[p_] != '"') fail("expected a string");
    std::string out;
    for (++p_; p_ < s_.size() && s_[p_] != '"'; ++p_) {
      if (s_[p_] == '\\' && p_ + 1 < s_.size()) ++p_;
      out.push_back(s_[p_]);
    }
    if (p_ >= s_.size()) fail("unterminated string");
    ++p_;
    return out;
  }
  node value() {
    ws();
    node n;
    if (p_ >= s_.size()) fail("unexpected end");
    if (s_[p_] == '[') {
      ++p_;
      n.leaf = false;
      if (!eat(']')) {
        do n.items.push_back(value());
        while (eat(','));
        if (!eat(']')) fail("expected ]");
      }
      return n;
    }
    if (s_[p_] == '"') {
      const std::string t = str();
      n.integer = false;
      const double inf = std::numeric_limits<double>::infinity();
      if (t == "NaN") n.value = std::numeric_limits<double>::quiet_NaN();
      else if (t == "Inf" || t == "Infinity" || t == "+Inf" || t == "+Infinity") n.value = inf;
      else if (t == "-Inf" || t == "-Infinity") n.value = -inf;
      else fail("string value \"" + t + "\"");
      return n;
    }
    const size_t b = p_;
    while (p_ < s_.size() && (std::isdigit(static_cast<unsigned char>(s_[p_])) || s_[p_] == '-' ||
                              s_[p_] == '+' || s_[p_] == '.' || s_[p_] == 'e' || s_[p_] == 'E'))
      ++p_;
    if (p_ == b) fail("expected a number");
    const std::string t = s_.substr(b, p_ - b);
    n.integer = t.find_first_of(".eE") == std::string::npos;
    char* end = nullptr;
    n.value = std::strtod(t.c_str(), &end);
    if (end != t.c_str() + t.size()) fail("bad number " + t);
    return n;
  }
  // dimensions of a rectangular array (throws if ragged)
  static void shape(const node& n, size_t depth, std::vector<size_t>& dims) {
    if (n.leaf) {
      if (depth != dims.size()) throw std::runtime_error("JSON data: ragged array");
      return;
    }
    if (depth == dims.size()) dims.push_back(n.items.size());
    else if (dims[depth] != n.items.size()) throw std::runtime_error("JSON data: ragged array");
    for (const auto& c : n.items) shape(c, depth + 1, dims);
  }
  // element at a multi-index
  static const node& at(const node& n, const std::vector<size_t>& idx) {
    const node* c = &n;
    for (size_t i : idx) c = &c->items[i];
    return *c;
  }
  void parse() {
    if (!eat('{')) fail("expected {");
    if (eat('}')) return;
    do {
      const std::string name = str();
      if (!eat(':')) fail("expected :");
      const node n = value();
      std::vector<size_t> dims;
      if (!n.leaf) {
        shape(n, 0, dims);
        // the innermost empty arrays give no leaf: dims stop at the zero
      }
      size_t total = 1;
      for (size_t d : dims) total *= d;
      std::vector<double> vals(total);
      bool integer = true;
      // column-major: first index fastest
      std::vector<size_t> idx(dims.size(), 0);
      for (size_t k = 0; k < total; ++k) {
        size_t r = k;
        for (size_t d = 0; d < dims.size(); ++d) {
          idx[d] = r % dims[d];
          r /= dims[d];
        }
        const node& leaf = at(n, idx);
        if (!leaf.leaf) throw std::runtime_error("JSON data: ragged array in " + name);
        vals[k] = leaf.value;
        integer = integer && leaf.integer;
      }
      add(name, vals, dims);
      names_.push_back(name);
      if (integer) ints_.push_back(name);
    } while (eat(','));
    if (!eat('}')) fail("expected }");
  }

  std::string s_;
  size_t p_ = 0;
  std::vector<std::string> names_, ints_;
};

}  // namespace io
}  // namespace stanx
#endif
