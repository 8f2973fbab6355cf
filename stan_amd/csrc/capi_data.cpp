// C-ABI for data files: a GLM context created from a Stan JSON data file,
// read the way a generated model's constructor reads its var_context
// (src/stan/model/model_base.hpp:36-56, src/stan/io/json/json_data.hpp).
#include <algorithm>
#include <cstring>
#include <exception>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/stanx.h"
#include "../include/stanx/json_data.hpp"

namespace sx {
void set_error(const std::string& msg);
}

namespace {

const char* first_present(const stanx::io::json_data& d, std::initializer_list<const char*> names) {
  for (const char* n : names)
    if (d.contains_r(n)) return n;
  return nullptr;
}

}  // namespace

extern "C" {

int sx_json_var(const char* path, const char* name, double* vals, int64_t cap, int64_t* dims,
                int max_dims, int* ndims, int64_t* size) {
  if (!path || !name) {
    sx::set_error("sx_json_var: NULL argument");
    return SX_EINVAL;
  }
  try {
    const auto d = stanx::io::json_data::from_file(path);
    if (!d.contains_r(name)) {
      sx::set_error(std::string("sx_json_var: no variable ") + name);
      return SX_EINVAL;
    }
    const auto v = d.vals_r(name);
    const auto dm = d.dims_r(name);
    if (size) *size = static_cast<int64_t>(v.size());
    if (ndims) *ndims = static_cast<int>(dm.size());
    if (dims)
      for (int i = 0; i < max_dims && i < static_cast<int>(dm.size()); ++i) dims[i] = dm[i];
    if (vals) std::memcpy(vals, v.data(), sizeof(double) * std::min<size_t>(v.size(), cap));
    return SX_OK;
  } catch (const std::exception& e) {
    sx::set_error(e.what());
    return SX_EINVAL;
  }
}

int sx_glm_create_json(sx_glm** out, const char* path, int family, int precision,
                       const sx_prior* prior, int device) {
  if (!out || !path) {
    sx::set_error("sx_glm_create_json: NULL argument");
    return SX_EINVAL;
  }
  *out = nullptr;
  std::vector<double> X, yd;
  std::vector<int32_t> yi;
  std::vector<int64_t> offsets;
  int64_t N = 0, K = 0, J = 0;
  try {
    const auto d = stanx::io::json_data::from_file(path);
    const char* xn = first_present(d, {"X", "x"});
    const char* yn = first_present(d, {"y", "Y"});
    if (!xn || !yn) {
      sx::set_error("sx_glm_create_json: the data need X and y (or Y)");
      return SX_EINVAL;
    }
    const auto xd = d.dims_r(xn);
    if (xd.size() != 2) {
      sx::set_error("sx_glm_create_json: X must be an N x K matrix");
      return SX_EINVAL;
    }
    N = static_cast<int64_t>(xd[0]);
    K = static_cast<int64_t>(xd[1]);
    if (d.contains_r("N") && static_cast<int64_t>(d.vals_r("N").at(0)) != N) {
      sx::set_error("sx_glm_create_json: N does not match the rows of X");
      return SX_EINVAL;
    }
    if (d.contains_r("K") && static_cast<int64_t>(d.vals_r("K").at(0)) != K) {
      sx::set_error("sx_glm_create_json: K does not match the columns of X");
      return SX_EINVAL;
    }
    const auto xc = d.vals_r(xn);  // column-major
    const auto yv = d.vals_r(yn);
    if (static_cast<int64_t>(yv.size()) != N) {
      sx::set_error("sx_glm_create_json: y must have N values");
      return SX_EINVAL;
    }
    // row order: as given, or grouped by the 1-based group index
    std::vector<int64_t> order(static_cast<size_t>(N));
    std::iota(order.begin(), order.end(), int64_t(0));
    if (family == SX_HIER_LOGIT) {
      const char* gn = first_present(d, {"group", "ids", "g"});
      if (!gn) {
        sx::set_error("sx_glm_create_json: the hierarchical model needs a group index (group or ids)");
        return SX_EINVAL;
      }
      const auto g = d.vals_r(gn);
      if (static_cast<int64_t>(g.size()) != N) {
        sx::set_error("sx_glm_create_json: the group index must have N values");
        return SX_EINVAL;
      }
      J = d.contains_r("J") ? static_cast<int64_t>(d.vals_r("J").at(0))
                            : (N ? static_cast<int64_t>(*std::max_element(g.begin(), g.end())) : 0);
      for (double v : g)
        if (v < 1 || v > static_cast<double>(J) || v != std::floor(v)) {
          sx::set_error("sx_glm_create_json: group indices must be integers in 1..J");
          return SX_EINVAL;
        }
      std::stable_sort(order.begin(), order.end(),
                       [&](int64_t a, int64_t b) { return g[a] < g[b]; });
      offsets.assign(static_cast<size_t>(J + 1), 0);
      for (double v : g) ++offsets[static_cast<size_t>(v)];
      for (int64_t j = 0; j < J; ++j) offsets[j + 1] += offsets[j];
    }
    X.resize(static_cast<size_t>(N * K));
    for (int64_t r = 0; r < N; ++r)
      for (int64_t k = 0; k < K; ++k) X[r * K + k] = xc[k * N + order[r]];
    if (family == SX_NORMAL) {
      yd.resize(static_cast<size_t>(N));
      for (int64_t r = 0; r < N; ++r) yd[r] = yv[order[r]];
    } else {
      yi.resize(static_cast<size_t>(N));
      for (int64_t r = 0; r < N; ++r) yi[r] = static_cast<int32_t>(yv[order[r]]);
    }
  } catch (const std::exception& e) {
    sx::set_error(e.what());
    return SX_EINVAL;
  }
  const void* y = family == SX_NORMAL ? static_cast<const void*>(yd.data())
                                      : static_cast<const void*>(yi.data());
  return sx_glm_create(out, family, X.data(), y, N, K, offsets.empty() ? nullptr : offsets.data(),
                       J, precision, prior, device);
}

}  // extern "C"
