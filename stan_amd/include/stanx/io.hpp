// Minimal data/init containers mirroring src/stan/io/var_context.hpp:30-140
// and array_var_context: named real arrays with dimensions.
#ifndef STANX_IO_HPP
#define STANX_IO_HPP

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace stanx {
namespace io {

class var_context {
 public:
  virtual ~var_context() {}
  virtual bool contains_r(const std::string& name) const = 0;
  virtual std::vector<double> vals_r(const std::string& name) const = 0;
  virtual std::vector<size_t> dims_r(const std::string& name) const = 0;
};

class array_var_context : public var_context {
 public:
  array_var_context() {}
  void add(const std::string& name, const std::vector<double>& vals,
           const std::vector<size_t>& dims) {
    vals_[name] = vals;
    dims_[name] = dims;
  }
  bool contains_r(const std::string& name) const override {
    return vals_.count(name) > 0;
  }
  std::vector<double> vals_r(const std::string& name) const override {
    auto it = vals_.find(name);
    if (it == vals_.end()) throw std::out_of_range("no variable " + name);
    return it->second;
  }
  std::vector<size_t> dims_r(const std::string& name) const override {
    auto it = dims_.find(name);
    if (it == dims_.end()) throw std::out_of_range("no variable " + name);
    return it->second;
  }

 private:
  std::map<std::string, std::vector<double>> vals_;
  std::map<std::string, std::vector<size_t>> dims_;
};

class empty_var_context : public array_var_context {};

}  // namespace io
}  // namespace stanx
#endif
