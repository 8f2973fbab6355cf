// C-ABI for the GLM log density + gradient engine (declared in
// include/stanx.h).  Owns the device copies of X and y in the tiled layout,
// a private HIP stream, pinned staging buffers and the kernel choice.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sx_internal.hpp"

namespace sx {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

}  // namespace sx

using sx::set_error;

namespace {

int choose_geom(sx_glm* m, int L_req, int smax_req) {
  int L = L_req, smax = smax_req;
  if (L <= 0) {
    // widest register tile that still fits: fewest lanes per row with
    // ceil(K / L) <= 32
    L = 1;
    while (L < 64 && (m->K + L - 1) / L > 32) L <<= 1;
  }
  const int64_t S = (m->K + L - 1) / L;
  if (smax <= 0) {
    smax = 4;
    while (smax < S) smax <<= 1;
  }
  if (S > smax || smax > 32 || (L & (L - 1)) != 0 || L > 64) {
    set_error("unsupported tile configuration for K=" + std::to_string(m->K));
    return SX_EINVAL;
  }
  m->geom = sx::make_geom(m->N, m->K, L, smax);
  return SX_OK;
}

int set_blocks(sx_glm* m, int blocks_req) {
  int dev = 0;
  SX_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  SX_HIP(hipGetDeviceProperties(&prop, dev));
  const int per_cu = sx::max_blocks_per_cu(m->family, m->geom.smax, m->threads);
  int blocks = blocks_req > 0 ? blocks_req : prop.multiProcessorCount * per_cu;
  const int64_t waves_needed = m->geom.T;  // one tile per wave at least
  const int64_t max_blocks = std::max<int64_t>(1, (waves_needed + 3) / 4);
  blocks = static_cast<int>(std::min<int64_t>(blocks, max_blocks));
  m->blocks = std::max(1, blocks);
  return SX_OK;
}

int alloc_work(sx_glm* m) {
  const int E = static_cast<int>(m->K) + 3;
  if (m->d_partials) (void)hipFree(m->d_partials);
  SX_HIP(hipMalloc(&m->d_partials, sizeof(double) * static_cast<size_t>(m->blocks) * E));
  return SX_OK;
}

}  // namespace

extern "C" {

const char* sx_last_error(void) { return sx::g_last_error.c_str(); }

int sx_glm_create(sx_glm** out, int family, const double* X, const void* y, int64_t N,
                  int64_t K, const int64_t* group_offsets, int64_t J, int precision,
                  const sx_prior* prior, int device) {
  if (!out) {
    set_error("sx_glm_create: out is NULL");
    return SX_EINVAL;
  }
  *out = nullptr;
  if (family != SX_LOGIT && family != SX_NORMAL) {
    set_error("sx_glm_create: family not supported by this build");
    return SX_EINVAL;
  }
  if (precision != SX_FP64) {
    set_error("sx_glm_create: precision not supported for this family");
    return SX_EINVAL;
  }
  if (N < 0 || K < 0 || K > 1024 || (N > 0 && (!X || !y))) {
    set_error("sx_glm_create: invalid sizes or NULL data");
    return SX_EINVAL;
  }
  (void)group_offsets;
  (void)J;
  // data checks the reference's GLM functions perform (y in {0,1},
  // finite y and X) happen once here instead of per evaluation
  if (family == SX_LOGIT) {
    const int32_t* yi = static_cast<const int32_t*>(y);
    for (int64_t n = 0; n < N; ++n)
      if (yi[n] != 0 && yi[n] != 1) {
        set_error("bernoulli_logit_glm: y[" + std::to_string(n) + "] is " +
                  std::to_string(yi[n]) + ", but must be in {0, 1}");
        return SX_EINVAL;
      }
  } else {
    const double* yd = static_cast<const double*>(y);
    for (int64_t n = 0; n < N; ++n)
      if (!std::isfinite(yd[n])) {
        set_error("normal_id_glm: y[" + std::to_string(n) + "] is not finite");
        return SX_EINVAL;
      }
  }
  for (int64_t i = 0; i < N * K; ++i)
    if (!std::isfinite(X[i])) {
      set_error("GLM: X contains a non-finite value");
      return SX_EINVAL;
    }
  int ndev = 0;
  SX_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) {
    set_error("sx_glm_create: no such device");
    return SX_EINVAL;
  }
  SX_HIP(hipSetDevice(device));
  sx_glm* m = new sx_glm();
  m->family = family;
  m->precision = precision;
  m->device = device;
  m->N = N;
  m->K = K;
  m->J = J;
  m->D = family == SX_LOGIT ? K + 1 : K + 2;
  if (prior) m->prior = *prior;
  else m->prior = sx_prior{5.0, 2.0, 2.0, 1.0};
  int rc = choose_geom(m, 0, 0);
  if (rc) {
    delete m;
    return rc;
  }
  auto fail = [&](int code) {
    sx_glm_destroy(m);
    return code;
  };
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    delete m;
    return SX_EDEVICE;
  }
  const sx::tile_geom& g = m->geom;
  const size_t xt_elems = static_cast<size_t>(g.T) * g.S * 64;
  if (xt_elems && hipMalloc(&m->d_X, xt_elems * sizeof(double)) != hipSuccess) {
    set_error("hipMalloc(X) failed");
    return fail(SX_ENOMEM);
  }
  if (xt_elems) {
    double* tmp = nullptr;
    if (hipMalloc(&tmp, static_cast<size_t>(N * K) * sizeof(double)) != hipSuccess) {
      set_error("hipMalloc(staging X) failed");
      return fail(SX_ENOMEM);
    }
    hipError_t e = hipMemcpy(tmp, X, static_cast<size_t>(N * K) * sizeof(double),
                             hipMemcpyHostToDevice);
    if (e == hipSuccess) e = sx::launch_tile_x_f64(tmp, m->d_X, N, K, g, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) {
      set_error(std::string("X upload: ") + hipGetErrorString(e));
      return fail(SX_EDEVICE);
    }
  }
  if (family == SX_LOGIT) {
    std::vector<int8_t> y8(static_cast<size_t>(std::max<int64_t>(N, 1)));
    const int32_t* yi = static_cast<const int32_t*>(y);
    for (int64_t n = 0; n < N; ++n) y8[n] = static_cast<int8_t>(yi[n]);
    if (hipMalloc(&m->d_y, y8.size()) != hipSuccess) return fail(SX_ENOMEM);
    if (hipMemcpy(m->d_y, y8.data(), y8.size(), hipMemcpyHostToDevice) != hipSuccess)
      return fail(SX_EDEVICE);
  } else {
    const size_t bytes = sizeof(double) * static_cast<size_t>(std::max<int64_t>(N, 1));
    if (hipMalloc(&m->d_y, bytes) != hipSuccess) return fail(SX_ENOMEM);
    if (N && hipMemcpy(m->d_y, y, sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess)
      return fail(SX_EDEVICE);
  }
  if (hipMalloc(&m->d_theta, sizeof(double) * (m->D + 1)) != hipSuccess) return fail(SX_ENOMEM);
  if (hipMalloc(&m->d_out, sizeof(double) * (m->D + 2)) != hipSuccess) return fail(SX_ENOMEM);
  if (hipMalloc(&m->d_counter, sizeof(unsigned)) != hipSuccess) return fail(SX_ENOMEM);
  if (hipMemset(m->d_counter, 0, sizeof(unsigned)) != hipSuccess) return fail(SX_EDEVICE);
  if (hipHostMalloc(&m->h_theta, sizeof(double) * (m->D + 1)) != hipSuccess) return fail(SX_ENOMEM);
  if (hipHostMalloc(&m->h_out, sizeof(double) * (m->D + 2)) != hipSuccess) return fail(SX_ENOMEM);
  if (hipEventCreate(&m->ev0) != hipSuccess || hipEventCreate(&m->ev1) != hipSuccess)
    return fail(SX_EDEVICE);
  if ((rc = set_blocks(m, 0)) || (rc = alloc_work(m))) return fail(rc);
  *out = m;
  return SX_OK;
}

void sx_glm_destroy(sx_glm* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  (void)hipFree(m->d_X);
  (void)hipFree(m->d_y);
  (void)hipFree(m->d_theta);
  (void)hipFree(m->d_partials);
  (void)hipFree(m->d_counter);
  (void)hipFree(m->d_out);
  if (m->h_theta) (void)hipHostFree(m->h_theta);
  if (m->h_out) (void)hipHostFree(m->h_out);
  if (m->ev0) (void)hipEventDestroy(m->ev0);
  if (m->ev1) (void)hipEventDestroy(m->ev1);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int sx_glm_num_params(const sx_glm* m, int64_t* D) {
  if (!m || !D) {
    set_error("sx_glm_num_params: NULL argument");
    return SX_EINVAL;
  }
  *D = m->D;
  return SX_OK;
}

int sx_glm_family(const sx_glm* m) { return m ? m->family : -1; }

}  // extern "C"

namespace {

int enqueue_pass(sx_glm* m, int propto, int jacobian) {
  sx::pass_args a;
  a.X = m->d_X;
  a.y = m->d_y;
  a.theta = m->d_theta;
  a.partials = m->d_partials;
  a.counter = m->d_counter;
  a.out = m->d_out;
  a.N = m->N;
  a.T = m->geom.T;
  a.K = static_cast<int>(m->K);
  a.L = m->geom.L;
  a.lgL = m->geom.lgL;
  a.S = m->geom.S;
  a.D = static_cast<int>(m->D);
  a.E = static_cast<int>(m->K) + 3;
  a.propto = propto;
  a.jacobian = jacobian;
  a.alpha_sd = m->prior.alpha_sd;
  a.beta_sd = m->prior.beta_sd;
  a.sigma_sd = m->prior.sigma_sd;
  SX_HIP(sx::launch_pass_f64(m->family, m->geom, a, m->blocks, m->threads, m->stream));
  return SX_OK;
}

int eval(sx_glm* m, const double* theta, int propto, int jacobian, double* lp,
         double* grad) {
  if (!m || !theta || !lp) {
    set_error("NULL argument");
    return SX_EINVAL;
  }
  SX_HIP(hipSetDevice(m->device));
  const size_t D = static_cast<size_t>(m->D);
  for (size_t i = 0; i < D; ++i)
    if (!std::isfinite(theta[i])) {
      set_error("log_prob: parameter " + std::to_string(i) + " is not finite");
      return SX_EDOMAIN;
    }
  std::memcpy(m->h_theta, theta, D * sizeof(double));
  SX_HIP(hipMemcpyAsync(m->d_theta, m->h_theta, D * sizeof(double), hipMemcpyHostToDevice,
                        m->stream));
  int rc = enqueue_pass(m, propto, jacobian);
  if (rc) return rc;
  SX_HIP(hipMemcpyAsync(m->h_out, m->d_out, (D + 2) * sizeof(double), hipMemcpyDeviceToHost,
                        m->stream));
  SX_HIP(hipStreamSynchronize(m->stream));
  *lp = m->h_out[0];
  if (grad) std::memcpy(grad, m->h_out + 1, D * sizeof(double));
  if (m->h_out[D + 1] != 0.0) {
    set_error(m->family == SX_LOGIT
                  ? "bernoulli_logit_glm_lpmf: linear predictor or log density is not finite"
                  : "normal_id_glm_lpdf: scale, linear predictor or log density is not finite");
    return SX_EDOMAIN;
  }
  return SX_OK;
}

}  // namespace

extern "C" {

int sx_glm_logp_grad(sx_glm* m, const double* theta, int propto, int jacobian, double* lp,
                     double* grad) {
  return eval(m, theta, propto, jacobian, lp, grad);
}

int sx_glm_logp(sx_glm* m, const double* theta, int propto, int jacobian, double* lp) {
  return eval(m, theta, propto, jacobian, lp, nullptr);
}

int sx_glm_logp_grad_batch(sx_glm* m, const double* theta, int64_t C, int propto,
                           int jacobian, double* lp, double* grad) {
  if (!m || C < 0 || (C > 0 && (!theta || !lp))) {
    set_error("sx_glm_logp_grad_batch: invalid argument");
    return SX_EINVAL;
  }
  int worst = SX_OK;
  for (int64_t c = 0; c < C; ++c) {
    int rc = eval(m, theta + c * m->D, propto, jacobian, lp + c,
                  grad ? grad + c * m->D : nullptr);
    if (rc == SX_EDOMAIN) {
      lp[c] = -INFINITY;
      worst = SX_EDOMAIN;
    } else if (rc) {
      return rc;
    }
  }
  return worst;
}

int sx_glm_time_logp_grad(sx_glm* m, const double* theta, int64_t C, int iters, int warmup,
                          double* ms_per_eval, double* ms_kernel) {
  if (!m || !theta || iters <= 0 || C != 1) {
    set_error("sx_glm_time_logp_grad: invalid argument");
    return SX_EINVAL;
  }
  SX_HIP(hipSetDevice(m->device));
  const size_t D = static_cast<size_t>(m->D);
  SX_HIP(hipMemcpy(m->d_theta, theta, D * sizeof(double), hipMemcpyHostToDevice));
  for (int i = 0; i < warmup; ++i) {
    int rc = enqueue_pass(m, 1, 1);
    if (rc) return rc;
  }
  SX_HIP(hipStreamSynchronize(m->stream));
  SX_HIP(hipEventRecord(m->ev0, m->stream));
  for (int i = 0; i < iters; ++i) {
    int rc = enqueue_pass(m, 1, 1);
    if (rc) return rc;
  }
  SX_HIP(hipEventRecord(m->ev1, m->stream));
  SX_HIP(hipEventSynchronize(m->ev1));
  float ms = 0.f;
  SX_HIP(hipEventElapsedTime(&ms, m->ev0, m->ev1));
  // the pass is a single kernel: per-launch time is the kernel time
  if (ms_per_eval) *ms_per_eval = ms / iters;
  if (ms_kernel) *ms_kernel = ms / iters;
  return SX_OK;
}

int sx_glm_kernel_config(const sx_glm* m, int* lanes_per_row, int* slots_per_lane,
                         int* blocks, int* threads) {
  if (!m) {
    set_error("NULL model");
    return SX_EINVAL;
  }
  if (lanes_per_row) *lanes_per_row = m->geom.L;
  if (slots_per_lane) *slots_per_lane = m->geom.S;
  if (blocks) *blocks = m->blocks;
  if (threads) *threads = m->threads;
  return SX_OK;
}

}  // extern "C"
