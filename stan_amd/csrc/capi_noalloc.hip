// The allocation-free entry point: one log density + gradient evaluation at
// a device-resident theta into a device-resident output, enqueued on a
// caller's stream, with every intermediate in a fixed per-device scratch
// array instead of library- or caller-allocated workspace.  For callers
// that may not allocate device memory (code under HIP graph capture, tight
// serving loops).
//
// The array is a file-scope __device__ object: each GPU the library runs on
// holds its own copy, at its own address, until the process ends.  Host
// code never passes the array's name to a kernel (on the host the name is
// only a stand-in for the symbol); the device address is looked up once
// per device with hipGetSymbolAddress when a context is created there, and
// that pointer is what the kernels receive.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>

#include "sx_internal.hpp"

namespace {

// 8 MiB of FP64 partial sums per GPU
constexpr int64_t kScratchDoubles = SX_NOALLOC_SCRATCH_BYTES / 8;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

}  // namespace

namespace sx {

int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("device index out of range");
    return SX_EINVAL;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch[device]) return SX_OK;
  int prev = 0;
  SX_HIP(hipGetDevice(&prev));
  SX_HIP(hipSetDevice(device));
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  (void)hipSetDevice(prev);
  if (e != hipSuccess || !p) {
    set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return SX_EDEVICE;
  }
  g_scratch[device] = static_cast<double*>(p);
  return SX_OK;
}

}  // namespace sx

extern "C" {

int sx_glm_logp_grad_noalloc(sx_glm* m, const double* d_theta, int propto, int jacobian,
                             double* d_out, void* stream) {
  if (!m || !d_theta || !d_out) {
    sx::set_error("sx_glm_logp_grad_noalloc: NULL argument");
    return SX_EINVAL;
  }
  double* base = (m->device >= 0 && m->device < kMaxDevices) ? g_scratch[m->device] : nullptr;
  if (!base) {
    sx::set_error("sx_glm_logp_grad_noalloc: no scratch address for this device");
    return SX_EDEVICE;
  }
  SX_HIP(hipSetDevice(m->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : m->stream;
  if (m->family == SX_HIER_LOGIT) {
    // chunk partials, then group-block rows (one chain)
    const int64_t E = m->K + 3, G = 2 * m->K + 5;
    const int64_t nblk = (m->J + 7) / 8;
    const int64_t need = E * (m->n_chunks > 0 ? m->n_chunks : 1) + G * nblk;
    if (need > kScratchDoubles) {
      sx::set_error("sx_glm_logp_grad_noalloc: partial sums (" + std::to_string(need * 8) +
                    " bytes) exceed the scratch array (" +
                    std::to_string(SX_NOALLOC_SCRATCH_BYTES) + " bytes)");
      return SX_EINVAL;
    }
    sx::hier_workspace w;
    w.cap = 1;  // one chain: no allocation happens in hier_enqueue
    w.cpart = base;
    w.gpart = base + E * (m->n_chunks > 0 ? m->n_chunks : 1);
    return sx::hier_enqueue(m, w, d_theta, 1, d_out, propto, jacobian, false, s);
  }
  const int64_t need = static_cast<int64_t>(m->blocks) * (m->K + 3);
  if (need > kScratchDoubles) {
    sx::set_error("sx_glm_logp_grad_noalloc: partial sums (" + std::to_string(need * 8) +
                  " bytes) exceed the scratch array (" +
                  std::to_string(SX_NOALLOC_SCRATCH_BYTES) + " bytes)");
    return SX_EINVAL;
  }
  return sx::plain_enqueue(m, d_theta, d_out, base, propto, jacobian, s, true);
}

}  // extern "C"
