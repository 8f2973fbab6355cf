// A measured HBM ceiling for the roofline figures: a plain device-to-device
// copy with 16-byte loads and stores over a grid-stride loop, timed with
// HIP events.  It reads and writes `bytes` each; the achieved rate counts
// both directions (2 * bytes per launch).  No counterpart in the
// reference.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "sx_internal.hpp"

namespace sx {
namespace {

__global__ __launch_bounds__(256) void copy16(const uint4* __restrict__ src,
                                              uint4* __restrict__ dst, int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride)
    dst[i] = src[i];
}

}  // namespace
}  // namespace sx

extern "C" int sx_device_copy_bandwidth(int device, int64_t bytes, int iters, double* gbs,
                                        double* ms_per_copy) {
  if (bytes < 16 || iters < 1 || !gbs) {
    sx::set_error("sx_device_copy_bandwidth: invalid argument");
    return SX_EINVAL;
  }
  SX_HIP(hipSetDevice(device));
  int cus = 0;
  SX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  const int64_t n = bytes / 16;
  void *src = nullptr, *dst = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = SX_OK;
  auto check = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == SX_OK) {
      sx::set_error(std::string("sx_device_copy_bandwidth: ") + what + ": " +
                    hipGetErrorString(e));
      rc = SX_EDEVICE;
    }
    return rc == SX_OK;
  };
  if (check(hipMalloc(&src, n * 16), "hipMalloc") && check(hipMalloc(&dst, n * 16), "hipMalloc") &&
      check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream") &&
      check(hipEventCreate(&e0), "event") && check(hipEventCreate(&e1), "event") &&
      check(hipMemsetAsync(src, 1, n * 16, s), "memset")) {
    const int blocks = 8 * cus;
    hipLaunchKernelGGL(sx::copy16, dim3(blocks), dim3(256), 0, s,
                       static_cast<const uint4*>(src), static_cast<uint4*>(dst), n);  // warm-up
    check(hipGetLastError(), "launch");
    check(hipEventRecord(e0, s), "event record");
    for (int i = 0; i < iters && rc == SX_OK; ++i) {
      hipLaunchKernelGGL(sx::copy16, dim3(blocks), dim3(256), 0, s,
                         static_cast<const uint4*>(src), static_cast<uint4*>(dst), n);
      check(hipGetLastError(), "launch");
    }
    check(hipEventRecord(e1, s), "event record");
    if (check(hipEventSynchronize(e1), "sync")) {
      float ms = 0.f;
      if (check(hipEventElapsedTime(&ms, e0, e1), "elapsed")) {
        const double per = static_cast<double>(ms) / iters;
        *gbs = 2.0 * static_cast<double>(n) * 16.0 / (per * 1e-3) / 1e9;
        if (ms_per_copy) *ms_per_copy = per;
      }
    }
  }
  if (s) (void)hipStreamSynchronize(s);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  return rc;
}
