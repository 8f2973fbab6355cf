// Output and control callbacks of the sampler services, mirroring the
// reference's interfaces (src/stan/callbacks/writer.hpp:16-61,
// logger.hpp:24-97, interrupt.hpp:14-28, stream_writer.hpp:14-93,
// structured_writer.hpp:16-170).
#ifndef STANX_CALLBACKS_HPP
#define STANX_CALLBACKS_HPP

#include <cmath>
#include <cstdio>
#include <ostream>
#include <sstream>
#include <string>
#include <vector>

namespace stanx {
namespace callbacks {

class writer {
 public:
  virtual ~writer() {}
  virtual void operator()(const std::vector<std::string>& names) {}
  virtual void operator()(const std::vector<double>& state) {}
  virtual void operator()() {}
  virtual void operator()(const std::string& message) {}
};

// Key-value records (the metric file: step size, metric type, inverse
// metric).  The base class ignores everything.
class structured_writer {
 public:
  virtual ~structured_writer() {}
  virtual void begin_record() {}
  virtual void end_record() {}
  virtual void write(const std::string& key, const std::string& value) {}
  virtual void write(const std::string& key, double value) {}
  virtual void write(const std::string& key, const std::vector<double>& values) {}
  // a row-major rows x cols matrix
  virtual void write(const std::string& key, const std::vector<double>& values, size_t rows,
                     size_t cols) {}
};

// Writes each record as one JSON object: {"key": value, ...}.  Numbers
// carry 17 significant digits (a double survives the round trip);
// non-finite values are written as the strings "NaN", "Inf", "-Inf".
class json_writer : public structured_writer {
 public:
  explicit json_writer(std::ostream& out) : out_(out) {}
  void begin_record() override {
    out_ << "{";
    first_ = true;
  }
  void end_record() override { out_ << "\n}\n"; }
  void write(const std::string& key, const std::string& value) override {
    this->key(key);
    out_ << '"' << escape(value) << '"';
  }
  void write(const std::string& key, double value) override {
    this->key(key);
    number(value);
  }
  void write(const std::string& key, const std::vector<double>& values) override {
    this->key(key);
    row(values.data(), values.size());
  }
  void write(const std::string& key, const std::vector<double>& values, size_t rows,
             size_t cols) override {
    this->key(key);
    out_ << "[";
    for (size_t r = 0; r < rows; ++r) {
      if (r) out_ << ", ";
      row(values.data() + r * cols, cols);
    }
    out_ << "]";
  }

 private:
  void key(const std::string& k) {
    out_ << (first_ ? "\n  " : ",\n  ") << '"' << escape(k) << "\": ";
    first_ = false;
  }
  void row(const double* v, size_t n) {
    out_ << "[";
    for (size_t i = 0; i < n; ++i) {
      if (i) out_ << ", ";
      number(v[i]);
    }
    out_ << "]";
  }
  void number(double v) {
    if (std::isnan(v)) {
      out_ << "\"NaN\"";
    } else if (std::isinf(v)) {
      out_ << (v > 0 ? "\"Inf\"" : "\"-Inf\"");
    } else {
      char buf[32];
      std::snprintf(buf, sizeof buf, "%.17g", v);
      out_ << buf;
    }
  }
  static std::string escape(const std::string& s) {
    std::string o;
    for (char c : s) {
      if (c == '"' || c == '\\') o += '\\';
      o += c;
    }
    return o;
  }
  std::ostream& out_;
  bool first_ = true;
};

class logger {
 public:
  virtual ~logger() {}
  virtual void debug(const std::string& message) {}
  virtual void debug(const std::stringstream& message) {}
  virtual void info(const std::string& message) {}
  virtual void info(const std::stringstream& message) {}
  virtual void warn(const std::string& message) {}
  virtual void warn(const std::stringstream& message) {}
  virtual void error(const std::string& message) {}
  virtual void error(const std::stringstream& message) {}
  virtual void fatal(const std::string& message) {}
  virtual void fatal(const std::stringstream& message) {}
};

class interrupt {
 public:
  virtual ~interrupt() {}
  virtual void operator()() {}
};

// Writes names and values as comma-separated lines and messages as lines
// behind a comment prefix, with the stream's own number formatting: the
// sample file of the reference's services (a Stan CSV file with "# ").
class stream_writer : public writer {
 public:
  explicit stream_writer(std::ostream& output, const std::string& comment_prefix = "")
      : output_(output), comment_prefix_(comment_prefix) {}
  void operator()(const std::vector<std::string>& names) override { write_row(names); }
  void operator()(const std::vector<double>& state) override { write_row(state); }
  void operator()() override { output_ << comment_prefix_ << std::endl; }
  void operator()(const std::string& message) override {
    output_ << comment_prefix_ << message << std::endl;
  }

 private:
  template <class T>
  void write_row(const std::vector<T>& v) {
    if (v.empty()) return;
    for (size_t i = 0; i + 1 < v.size(); ++i) output_ << v[i] << ",";
    output_ << v.back() << std::endl;
  }
  std::ostream& output_;
  std::string comment_prefix_;
};

// Collects every numeric row and every message (tests and the C-ABI use it).
class table_writer : public writer {
 public:
  std::vector<std::string> names;
  std::vector<std::vector<double>> rows;
  std::vector<std::string> messages;
  void operator()(const std::vector<std::string>& n) override { names = n; }
  void operator()(const std::vector<double>& s) override { rows.push_back(s); }
  void operator()(const std::string& m) override { messages.push_back(m); }
};

class counting_logger : public logger {
 public:
  int n_info = 0, n_warn = 0, n_error = 0;
  std::vector<std::string> errors;
  void info(const std::string&) override { ++n_info; }
  void info(const std::stringstream&) override { ++n_info; }
  void warn(const std::string&) override { ++n_warn; }
  void warn(const std::stringstream&) override { ++n_warn; }
  void error(const std::string& m) override {
    ++n_error;
    errors.push_back(m);
  }
  void error(const std::stringstream& m) override {
    ++n_error;
    errors.push_back(m.str());
  }
};

}  // namespace callbacks
}  // namespace stanx
#endif
