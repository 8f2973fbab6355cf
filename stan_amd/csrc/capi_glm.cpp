// C-ABI for the GLM log density + gradient engine (declared in
// include/stanx.h).  Owns the device copies of X and y in the tiled layout,
// a private HIP stream, pinned staging buffers and the kernel choice.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
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

// compiled register capacities are the multiples of 4 up to 32
int smax_for(int64_t S) {
  int smax = 4;
  while (smax < S) smax += 4;
  return smax;
}

// lanes per row: fewest lanes whose register tile fits 32 slots
int default_lanes(int64_t K) {
  int L = 1;
  while (L < 64 && (K + L - 1) / L > 32) L <<= 1;
  return L;
}

int device_cus(int* cus) {
  int dev = 0;
  SX_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  SX_HIP(hipGetDeviceProperties(&prop, dev));
  *cus = prop.multiProcessorCount;
  return SX_OK;
}

int clamp_blocks(const sx_glm* m, int blocks) {
  const int64_t max_blocks = std::max<int64_t>(1, (m->geom.T + 3) / 4);
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(blocks, max_blocks)));
}

int retile(sx_glm* m, const double* d_rowmajor, int L, int smax, int prefetch) {
  sx::tile_geom g = sx::make_geom(m->N, m->K, L, smax);
  g.prefetch = prefetch;
  const size_t need = static_cast<size_t>(g.T) * g.S * 64;
  const size_t have = static_cast<size_t>(m->geom.T) * m->geom.S * 64;
  if (need != have || (!m->d_X && need)) {
    if (m->d_X) (void)hipFree(m->d_X);
    m->d_X = nullptr;
    if (need) SX_HIP(hipMalloc(&m->d_X, need * sizeof(double)));
  }
  if (need) SX_HIP(sx::launch_tile_x_f64(d_rowmajor, m->d_X, m->N, m->K, g, m->stream));
  m->geom = g;
  return SX_OK;
}

}  // namespace

// forward declarations used by the autotuner
namespace {
int enqueue_pass(sx_glm* m, int propto, int jacobian, bool reduce = true);
}

namespace {

// The autotuner's decisions, one per device model and data shape.  The
// layout (lanes per row) and the block count set the order of the FP64
// sums, so the gradient's last bits depend on the choice; timing alone
// could choose differently in two contexts of the same data.  The first
// decision for a (device, family, N, K) is therefore kept for the process
// and reused by every later context, which makes two autotuned contexts of
// one data set bitwise equal.  With STANX_TUNING_CACHE naming a file, the
// decisions are also read from and appended to it, so that separate
// processes (e.g. seeded reruns) share them too.
struct tuned {
  int L, smax, pf, blocks;
};
std::mutex g_tune_mu;
std::map<std::string, tuned> g_tuned;
bool g_tune_file_read = false;

std::string tune_key(int device, int family, int64_t N, int64_t K) {
  hipDeviceProp_t prop;
  std::string arch = "dev";
  if (hipGetDeviceProperties(&prop, device) == hipSuccess)
    arch = std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount);
  std::ostringstream k;
  k << arch << " " << family << " " << N << " " << K;
  return k.str();
}

// caller holds g_tune_mu
void read_tune_file() {
  if (g_tune_file_read) return;
  g_tune_file_read = true;
  const char* path = std::getenv("STANX_TUNING_CACHE");
  if (!path) return;
  std::ifstream f(path);
  std::string arch;
  int fam, L, smax, pf, blocks;
  long long N, K;
  while (f >> arch >> fam >> N >> K >> L >> smax >> pf >> blocks) {
    std::ostringstream k;
    k << arch << " " << fam << " " << N << " " << K;
    g_tuned.emplace(k.str(), tuned{L, smax, pf, blocks});
  }
}

void remember_tuning(const std::string& key, const tuned& t) {
  g_tuned.emplace(key, t);
  if (const char* path = std::getenv("STANX_TUNING_CACHE")) {
    std::ofstream f(path, std::ios::app);
    f << key << " " << t.L << " " << t.smax << " " << t.pf << " " << t.blocks << "\n";
  }
}

// Times candidate layouts / launch shapes on the device and keeps the
// fastest.  X is re-tiled from the row-major staging copy per layout.
int autotune(sx_glm* m, const double* d_rowmajor, int cus) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  read_tune_file();
  const std::string key = tune_key(m->device, m->family, m->N, m->K);
  const auto hit = g_tuned.find(key);
  if (hit != g_tuned.end()) {
    const tuned& t = hit->second;
    int rc = retile(m, d_rowmajor, t.L, t.smax, t.pf);
    if (rc) return rc;
    m->blocks = clamp_blocks(m, t.blocks);
    char line[160];
    std::snprintf(line, sizeof line,
                  "cached: L=%d smax=%d prefetch=%d blocks=%d (the first decision for %s)\n",
                  t.L, t.smax, t.pf, m->blocks, key.c_str());
    m->tuning_log = line;
    return SX_OK;
  }
  std::string log;
  std::vector<double> zeros(static_cast<size_t>(m->D), 0.0);
  SX_HIP(hipMemcpy(m->d_theta, zeros.data(), zeros.size() * sizeof(double),
                   hipMemcpyHostToDevice));
  struct cand { int L, smax, pf, blocks; double ms; };
  cand best{0, 0, 0, 0, 1e30};
  for (int L = 1; L <= 64; L <<= 1) {
    const int64_t S = (m->K + L - 1) / L;
    if (S > 32) continue;
    const double waste = static_cast<double>(S * L - m->K) / std::max<int64_t>(1, m->K);
    if (waste > 0.15 && L != default_lanes(m->K)) continue;
    const int smax = smax_for(S);
    for (int pf = 0; pf <= 1; ++pf) {
      int rc = retile(m, d_rowmajor, L, smax, pf);
      if (rc) return rc;
      for (int mult : {1, 2, 3, 4, 6, 8}) {
        m->blocks = clamp_blocks(m, mult * cus);
        if (mult > 1 && m->blocks < mult * cus) break;
        for (int w = 0; w < 2; ++w) {
          rc = enqueue_pass(m, 1, 1);
          if (rc) return rc;
        }
        SX_HIP(hipEventRecord(m->ev0, m->stream));
        const int reps = 5;
        for (int i = 0; i < reps; ++i) {
          rc = enqueue_pass(m, 1, 1);
          if (rc) return rc;
        }
        SX_HIP(hipEventRecord(m->ev1, m->stream));
        SX_HIP(hipEventSynchronize(m->ev1));
        float ms = 0.f;
        SX_HIP(hipEventElapsedTime(&ms, m->ev0, m->ev1));
        const double per = ms / reps;
        char line[160];
        std::snprintf(line, sizeof line, "L=%d S=%d smax=%d prefetch=%d blocks=%d: %.4f ms\n",
                      L, static_cast<int>(S), smax, pf, m->blocks, per);
        log += line;
        if (per < best.ms) best = cand{L, smax, pf, m->blocks, per};
      }
    }
  }
  int rc = retile(m, d_rowmajor, best.L, best.smax, best.pf);
  if (rc) return rc;
  m->blocks = best.blocks;
  remember_tuning(key, tuned{best.L, best.smax, best.pf, best.blocks});
  char line[160];
  std::snprintf(line, sizeof line, "chosen: L=%d smax=%d prefetch=%d blocks=%d (%.4f ms)\n",
                best.L, best.smax, best.pf, best.blocks, best.ms);
  log += line;
  m->tuning_log = log;
  return SX_OK;
}

// Hierarchical model: cut every group into chunks of at most `chunk` rows
// (a chunk never straddles a group) and upload the chunk tables.
int hier_chunks(sx_glm* m, const int64_t* off, int chunk) {
  std::vector<int> cg;
  std::vector<int64_t> cr0, gc0;
  gc0.reserve(static_cast<size_t>(m->J) + 1);
  for (int64_t j = 0; j < m->J; ++j) {
    gc0.push_back(static_cast<int64_t>(cg.size()));
    for (int64_t r = off[j]; r < off[j + 1]; r += chunk) {
      cg.push_back(static_cast<int>(j));
      cr0.push_back(r);
    }
  }
  gc0.push_back(static_cast<int64_t>(cg.size()));
  cr0.push_back(m->N);
  (void)hipFree(m->d_chunk_group);
  (void)hipFree(m->d_chunk_row0);
  (void)hipFree(m->d_group_chunk0);
  m->d_chunk_group = nullptr;
  m->d_chunk_row0 = m->d_group_chunk0 = nullptr;
  m->n_chunks = static_cast<int64_t>(cg.size());
  m->hier_chunk = chunk;
  SX_HIP(hipMalloc(&m->d_chunk_group, sizeof(int) * std::max<size_t>(1, cg.size())));
  SX_HIP(hipMalloc(&m->d_chunk_row0, sizeof(int64_t) * cr0.size()));
  SX_HIP(hipMalloc(&m->d_group_chunk0, sizeof(int64_t) * gc0.size()));
  if (!cg.empty())
    SX_HIP(hipMemcpy(m->d_chunk_group, cg.data(), sizeof(int) * cg.size(), hipMemcpyHostToDevice));
  SX_HIP(hipMemcpy(m->d_chunk_row0, cr0.data(), sizeof(int64_t) * cr0.size(), hipMemcpyHostToDevice));
  SX_HIP(hipMemcpy(m->d_group_chunk0, gc0.data(), sizeof(int64_t) * gc0.size(), hipMemcpyHostToDevice));
  return SX_OK;
}

// Chunk size and FP32 product kernel of the hierarchical model, timed on
// the device at context creation: 128, 256, 512 or 1024 rows, and for an
// FP32 context with K <= 31 the FP32 MFMA products or the BF16 three-term
// split products, each timed as full evaluations at theta = 0 (the
// kernels' work does not depend on the values) -- one chain in FP64 for an
// FP64 context, a batch of 64 chains for an FP32 one.  Both choices set
// the order (or rounding) of sums, so the first decision per (device,
// precision, N, J, K, forced kernel) is kept for the process and in
// STANX_TUNING_CACHE (family field 200 + [FP32] + 10 x forced kernel, K
// field 64 J + K).
int hier_autotune(sx_glm* m, const int64_t* off, int kernel) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  read_tune_file();
  const bool fp32 = m->precision == SX_FP32;
  const std::string key =
      tune_key(m->device, 200 + (fp32 ? 1 : 0) + 10 * kernel, m->N, 64 * m->J + m->K);
  const auto hit = g_tuned.find(key);
  if (hit != g_tuned.end()) {
    m->hier_x6 = hit->second.smax;
    char line[160];
    std::snprintf(line, sizeof line, "hierarchical chunk cached: %d rows, %s products\n",
                  hit->second.L, m->hier_x6 ? "BF16 split" : "FP32");
    m->tuning_log += line;
    return hier_chunks(m, off, hit->second.L);
  }
  const int C = fp32 ? 64 : 1;
  const int64_t D = m->D;
  double *d_th = nullptr, *d_out = nullptr;
  if (hipMalloc(&d_th, sizeof(double) * C * D) != hipSuccess ||
      hipMalloc(&d_out, sizeof(double) * C * (D + 2)) != hipSuccess) {
    (void)hipFree(d_th);
    set_error("hierarchical autotune: out of device memory");
    return SX_ENOMEM;
  }
  int rc = hipMemsetAsync(d_th, 0, sizeof(double) * C * D, m->stream) == hipSuccess
               ? SX_OK : SX_EDEVICE;
  std::string log;
  int best_c = sx::kHierChunk, best_k = 0;
  double best_ms = 1e30;
  const bool can_x6 = fp32 && m->K <= 31;
  const int k_lo = kernel == 2 && can_x6 ? 1 : 0;
  const int k_hi = kernel == 1 || !can_x6 ? 0 : 1;
  for (int cand = 0; cand < 4 * (k_hi - k_lo + 1); ++cand) {
    const int chunk = 128 << (cand / (k_hi - k_lo + 1));
    const int kx6 = k_lo + cand % (k_hi - k_lo + 1);
    if (rc) break;
    m->hier_x6 = kx6;
    rc = hier_chunks(m, off, chunk);
    if (rc) break;
    sx::hier_workspace w;
    for (int i = 0; i < 2 && rc == SX_OK; ++i)
      rc = sx::hier_enqueue(m, w, d_th, C, d_out, 1, 1, fp32, m->stream);
    float ms = 0.f;
    const int reps = 5;
    if (rc == SX_OK && hipEventRecord(m->ev0, m->stream) != hipSuccess) rc = SX_EDEVICE;
    for (int i = 0; i < reps && rc == SX_OK; ++i)
      rc = sx::hier_enqueue(m, w, d_th, C, d_out, 1, 1, fp32, m->stream);
    if (rc == SX_OK && (hipEventRecord(m->ev1, m->stream) != hipSuccess ||
                        hipEventSynchronize(m->ev1) != hipSuccess ||
                        hipEventElapsedTime(&ms, m->ev0, m->ev1) != hipSuccess))
      rc = SX_EDEVICE;
    (void)hipStreamSynchronize(m->stream);
    sx::free_workspace(w);
    if (rc) break;
    const double per = ms / reps;
    char line[160];
    std::snprintf(line, sizeof line,
                  "hierarchical chunk=%d rows (%lld chunks)%s: %.4f ms (%d chain%s)\n", chunk,
                  static_cast<long long>(m->n_chunks), fp32 ? (kx6 ? ", BF16 split products" : ", FP32 products") : "",
                  per, C, C > 1 ? "s" : "");
    log += line;
    if (per < best_ms) {
      best_ms = per;
      best_c = chunk;
      best_k = kx6;
    }
  }
  m->hier_x6 = best_k;
  (void)hipFree(d_th);
  (void)hipFree(d_out);
  m->tuning_log += log;
  if (rc) return rc;
  remember_tuning(key, tuned{best_c, best_k, 0, 0});
  char line[160];
  std::snprintf(line, sizeof line, "hierarchical chunk chosen: %d rows, %s products (%.4f ms)\n",
                best_c, best_k ? "BF16 split" : "FP32", best_ms);
  m->tuning_log += line;
  return hier_chunks(m, off, best_c);
}

// Hierarchical model: the chunk tables (the caller's chunk size, else the
// autotuned one, else kHierChunk), an FP32 copy of X for batched evaluations (SX_FP32), and the
// single-chain partials.  The FP32 copy is sign-folded: row n holds
// (2 y_n - 1) x_n.
int hier_setup(sx_glm* m, const double* d_X, const double* X, const int32_t* y,
               const int64_t* off, bool do_tune, int chunk, int kernel) {
  if (m->precision == SX_FP32 && m->N > 0) {
    // 32 K + 64 zero floats past the last row: the FP32 pass stages whole
    // 32-row tiles (64-float loads per lane) past a chunk's end
    const size_t n = static_cast<size_t>(m->N * m->K);
    const size_t K = static_cast<size_t>(m->K);
    const size_t total = n + 32 * K + 64;
    std::vector<float> x32(total, 0.0f);
    for (size_t i = 0; i < n; ++i) {
      const float v = static_cast<float>(X[i]);
      x32[i] = y[i / K] ? v : -v;
    }
    if (n > 0) {
      SX_HIP(hipMalloc(&m->d_X32, total * sizeof(float)));
      SX_HIP(hipMemcpy(m->d_X32, x32.data(), total * sizeof(float), hipMemcpyHostToDevice));
    }
    if (K <= 31) {
      // the x6 pass's copy: per row and term t, columns 0 .. 8 NCH - 1 of
      // (x'_0 .. x'_{K-1}, s, 0 ...) as term t of the exact three-term
      // BF16 split by truncation (h = x with the low 16 bits cleared,
      // m = (x - h) likewise, l = x - h - m), 32 zero rows past the last
      const size_t W = 8 * static_cast<size_t>(sx::hier_x6_chunks(static_cast<int>(K)));
      const size_t N = static_cast<size_t>(m->N);
      std::vector<uint16_t> xs((N + 32) * 3 * W, 0);
      auto bits = [](float f) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        return u;
      };
      auto from_bits = [](uint32_t u) {
        float f;
        std::memcpy(&f, &u, 4);
        return f;
      };
      for (size_t r = 0; r < N; ++r) {
        uint16_t* row = xs.data() + r * 3 * W;
        for (size_t c = 0; c <= K; ++c) {
          const float v = c < K ? x32[r * K + c] : (y[r] ? 1.0f : -1.0f);
          const uint32_t h = bits(v) & 0xffff0000u;
          const float r1 = v - from_bits(h);
          const uint32_t mm = bits(r1) & 0xffff0000u;
          const uint32_t l = bits(r1 - from_bits(mm));
          row[c] = static_cast<uint16_t>(h >> 16);
          row[W + c] = static_cast<uint16_t>(mm >> 16);
          row[2 * W + c] = static_cast<uint16_t>(l >> 16);
        }
      }
      SX_HIP(hipMalloc(&m->d_XS, xs.size() * sizeof(uint16_t)));
      SX_HIP(hipMemcpy(m->d_XS, xs.data(), xs.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    }
  }
  m->d_X = const_cast<double*>(d_X);  // the FP64 pass reads X row-major
  // FP32 products unless asked for (or timed) otherwise
  m->hier_x6 = kernel == 2 && m->precision == SX_FP32 && m->K <= 31 ? 1 : 0;
  const int rc = chunk > 0 ? hier_chunks(m, off, chunk)
                 : do_tune && m->N * m->K > 0 ? hier_autotune(m, off, kernel)
                                              : hier_chunks(m, off, sx::kHierChunk);
  if (rc) return rc;
  return sx::hier_prepare(m, m->hw, 1);
}

}  // namespace

extern "C" {

const char* sx_last_error(void) { return sx::g_last_error.c_str(); }

int sx_glm_create(sx_glm** out, int family, const double* X, const void* y, int64_t N,
                  int64_t K, const int64_t* group_offsets, int64_t J, int precision,
                  const sx_prior* prior, int device) {
  return sx_glm_create_ex(out, family, X, y, N, K, group_offsets, J, precision, prior, device,
                          nullptr);
}

int sx_glm_create_ex(sx_glm** out, int family, const double* X, const void* y, int64_t N,
                     int64_t K, const int64_t* group_offsets, int64_t J, int precision,
                     const sx_prior* prior, int device, const sx_tuning* tuning) {
  if (!out) {
    set_error("sx_glm_create: out is NULL");
    return SX_EINVAL;
  }
  *out = nullptr;
  if (family != SX_LOGIT && family != SX_NORMAL && family != SX_HIER_LOGIT) {
    set_error("sx_glm_create: unknown family");
    return SX_EINVAL;
  }
  const bool hier = family == SX_HIER_LOGIT;
  if ((!hier && precision != SX_FP64 && precision != SX_BF16X2) ||
      (hier && precision != SX_FP64 && precision != SX_FP32)) {
    set_error("sx_glm_create: precision not supported for this family");
    return SX_EINVAL;
  }
  if (hier) {
    if (K > 32) {
      set_error("sx_glm_create: the hierarchical model supports K <= 32");
      return SX_EINVAL;
    }
    if (J < 1 || !group_offsets) {
      set_error("sx_glm_create: the hierarchical model needs J >= 1 group offsets");
      return SX_EINVAL;
    }
    if (group_offsets[0] != 0 || group_offsets[J] != N) {
      set_error("sx_glm_create: group offsets must start at 0 and end at N");
      return SX_EINVAL;
    }
    for (int64_t j = 0; j < J; ++j)
      if (group_offsets[j + 1] < group_offsets[j]) {
        set_error("sx_glm_create: group offsets must be non-decreasing");
        return SX_EINVAL;
      }
  }
  if (precision == SX_BF16X2 && K > 512) {
    set_error("sx_glm_create: BF16 hi+lo batched evaluation supports K <= 512");
    return SX_EINVAL;
  }
  if (N < 0 || K < 0 || K > 1024 || (N > 0 && (!X || !y))) {
    set_error("sx_glm_create: invalid sizes or NULL data");
    return SX_EINVAL;
  }
  sx_tuning tune{0, 0, -1, 0, -1, 0, 0};
  if (tuning) tune = *tuning;
  if (tune.lanes_per_row < 0 || tune.lanes_per_row > 64 ||
      (tune.lanes_per_row & (tune.lanes_per_row - 1)) != 0 ||
      tune.slots_cap < 0 || tune.slots_cap > 32 || tune.slots_cap % 4 != 0 ||
      tune.prefetch < -1 || tune.prefetch > 1 || tune.blocks < 0 || tune.hier_chunk < 0 ||
      tune.hier_chunk > 4096 || tune.hier_chunk % 32 != 0 || tune.hier_kernel < 0 ||
      tune.hier_kernel > 2) {
    set_error("sx_glm_create: invalid tuning");
    return SX_EINVAL;
  }
  // data checks the reference's GLM functions perform (y in {0,1},
  // finite y and X) happen once here instead of per evaluation
  if (family == SX_LOGIT || hier) {
    const int32_t* yi = static_cast<const int32_t*>(y);
    for (int64_t n = 0; n < N; ++n)
      if (yi[n] != 0 && yi[n] != 1) {
        set_error(std::string(hier ? "bernoulli_logit" : "bernoulli_logit_glm") + ": y[" + std::to_string(n) + "] is " +
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
  int cus = 0;
  int rc = device_cus(&cus);
  if (rc) return rc;
  // the allocation-free entry point's scratch address on this device (its
  // failure only disables that entry point, which then reports it)
  (void)sx::noalloc_setup(device);
  sx_glm* m = new sx_glm();
  m->family = family;
  m->precision = precision;
  m->device = device;
  m->N = N;
  m->K = K;
  m->J = J;
  m->D = family == SX_LOGIT ? K + 1 : family == SX_NORMAL ? K + 2 : 2 * K + 2 + J * (K + 1);
  m->prior = prior ? *prior : sx_prior{5.0, 2.0, 2.0, 1.0};
  auto fail = [&](int code) {
    sx_glm_destroy(m);
    return code;
  };
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    delete m;
    return SX_EDEVICE;
  }
  if (family == SX_LOGIT || hier) {
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
  if (hipHostMalloc(&m->h_theta, sizeof(double) * (m->D + 1)) != hipSuccess) return fail(SX_ENOMEM);
  if (hipHostMalloc(&m->h_out, sizeof(double) * (m->D + 2)) != hipSuccess) return fail(SX_ENOMEM);
  if (hipEventCreate(&m->ev0) != hipSuccess || hipEventCreate(&m->ev1) != hipSuccess)
    return fail(SX_EDEVICE);
  const int max_blocks = 8 * cus;
  const size_t E = static_cast<size_t>(K) + 3;
  if (hipMalloc(&m->d_partials, sizeof(double) * max_blocks * E) != hipSuccess)
    return fail(SX_ENOMEM);
  m->max_blocks = max_blocks;

  // row-major staging copy on the device, tiled into the kernel layout (the
  // hierarchical model keeps it, with 64 K zero doubles past the last row:
  // its FP64 pass loads whole 64-row blocks past a chunk's end)
  double* staging = nullptr;
  if (N * K > 0) {
    const size_t pad = hier ? static_cast<size_t>(64 * K) : 0;
    if (hipMalloc(&staging, (static_cast<size_t>(N * K) + pad) * sizeof(double)) != hipSuccess) {
      set_error("hipMalloc(staging X) failed");
      return fail(SX_ENOMEM);
    }
    if (hipMemcpy(staging, X, static_cast<size_t>(N * K) * sizeof(double),
                  hipMemcpyHostToDevice) != hipSuccess ||
        (pad > 0 && hipMemset(staging + N * K, 0, pad * sizeof(double)) != hipSuccess)) {
      (void)hipFree(staging);
      set_error("X upload failed");
      return fail(SX_EDEVICE);
    }
  }
  const bool do_tune =
      tune.autotune == 1 || (tune.autotune == -1 && tuning == nullptr && N * K >= (1 << 22));
  if (hier) {
    rc = hier_setup(m, staging, X, static_cast<const int32_t*>(y), group_offsets, do_tune,
                    tune.hier_chunk, tune.hier_kernel);
    if (rc == SX_OK && hipStreamSynchronize(m->stream) != hipSuccess) rc = SX_EDEVICE;
    if (rc) {
      m->d_X = nullptr;
      if (staging) (void)hipFree(staging);
      return fail(rc);
    }
    *out = m;
    return SX_OK;
  }
  if (do_tune && N * K > 0) {
    rc = autotune(m, staging, cus);
  } else {
    const int L = tune.lanes_per_row > 0 ? tune.lanes_per_row : default_lanes(K);
    const int64_t S = (K + L - 1) / L;
    const int smax = tune.slots_cap > 0 ? tune.slots_cap : smax_for(S);
    if (S > smax) {
      set_error("sx_glm_create: K does not fit lanes_per_row x slots_cap");
      rc = SX_EINVAL;
    } else {
      rc = retile(m, staging, L, smax, tune.prefetch == 1 ? 1 : 0);
      m->blocks = clamp_blocks(m, tune.blocks > 0 ? std::min(tune.blocks, max_blocks)
                                                  : 2 * cus);
    }
  }
  if (rc == SX_OK && precision == SX_BF16X2) {
    m->Np = (N + 31) / 32 * 32;
    // a multiple of 32 up to 256; above, of 64 (the adjoint's column
    // blocks of 256 + 64 j)
    m->Kp = K <= 256 ? std::max<int64_t>(32, (K + 31) / 32 * 32) : (K + 63) / 64 * 64;
    const size_t elems = static_cast<size_t>(m->Np) * m->Kp;
    if (hipMalloc(&m->d_Xh, elems * 2) != hipSuccess || hipMalloc(&m->d_Xl, elems * 2) != hipSuccess ||
        hipMalloc(&m->d_yf, sizeof(float) * m->Np) != hipSuccess) {
      set_error("hipMalloc(BF16 X) failed");
      rc = SX_ENOMEM;
    } else if (sx::launch_split_x(staging, m->d_Xh, m->d_Xl, N, K, m->Np, m->Kp, m->stream) !=
                   hipSuccess ||
               sx::launch_prep_y(m->d_y, m->d_yf, family, N, m->Np, m->stream) != hipSuccess) {
      set_error("BF16 split of X failed");
      rc = SX_EDEVICE;
    } else if (do_tune && N * K > 0) {
      rc = sx::mx_autotune(m);
    }
  }
  if (rc == SX_OK && hipStreamSynchronize(m->stream) != hipSuccess) rc = SX_EDEVICE;
  if (staging) (void)hipFree(staging);
  if (rc) return fail(rc);
  *out = m;
  return SX_OK;
}

const char* sx_glm_tuning_log(const sx_glm* m) { return m ? m->tuning_log.c_str() : ""; }

void sx_glm_destroy(sx_glm* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  (void)hipFree(m->d_X);
  (void)hipFree(m->d_y);
  (void)hipFree(m->d_theta);
  (void)hipFree(m->d_partials);
  (void)hipFree(m->d_out);
  (void)hipFree(m->d_Xh);
  (void)hipFree(m->d_Xl);
  (void)hipFree(m->d_yf);
  sx::free_workspace(m->mx);
  (void)hipFree(m->d_btheta);
  (void)hipFree(m->d_bout);
  (void)hipFree(m->d_X32);
  (void)hipFree(m->d_XS);
  (void)hipFree(m->d_chunk_group);
  (void)hipFree(m->d_chunk_row0);
  (void)hipFree(m->d_group_chunk0);
  sx::free_workspace(m->hw);
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

int sx_glm_dims(const sx_glm* m, int64_t* N, int64_t* K, int64_t* J) {
  if (!m) {
    set_error("sx_glm_dims: NULL model");
    return SX_EINVAL;
  }
  if (N) *N = m->N;
  if (K) *K = m->K;
  if (J) *J = m->J;
  return SX_OK;
}

}  // extern "C"

namespace {

int enqueue_pass(sx_glm* m, int propto, int jacobian, bool reduce) {
  if (m->family == SX_HIER_LOGIT)
    return sx::hier_enqueue(m, m->hw, m->d_theta, 1, m->d_out, propto, jacobian, false, m->stream);
  return sx::plain_enqueue(m, m->d_theta, m->d_out, m->d_partials, propto, jacobian, m->stream,
                           reduce);
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
    set_error(m->family == SX_LOGIT ? "bernoulli_logit_glm_lpmf: linear predictor or log density is not finite"
              : m->family == SX_NORMAL
                  ? "normal_id_glm_lpdf: scale, linear predictor or log density is not finite"
                  : "bernoulli_logit_lpmf: linear predictor or log density is not finite");
    return SX_EDOMAIN;
  }
  return SX_OK;
}

}  // namespace

namespace sx {

// Batched MFMA evaluation of C chains: theta (device, C x D) -> out
// (device, C x (D + 2)).  Row splits depend on N only, so each chain's
// result is independent of the other chains in the batch.
void free_workspace(mx_workspace& w) {
  (void)hipFree(w.Bh);
  (void)hipFree(w.Bl);
  (void)hipFree(w.Bm);
  (void)hipFree(w.Rh);
  (void)hipFree(w.Rl);
  (void)hipFree(w.alpha);
  (void)hipFree(w.gpart);
  (void)hipFree(w.spart);
  w = mx_workspace();
}

void free_workspace(hier_workspace& w) {
  (void)hipFree(w.cpart);
  (void)hipFree(w.gpart);
  (void)hipFree(w.wsplit);
  w = hier_workspace();
}

// chains padded to the chain tile of the MFMA kernel (32x32 tiles, waves
// of 32 chains): 64 for batches of at most 64 chains (two waves per
// workgroup), else a multiple of 128 (four waves)
// The FP64 pass of a logistic or normal context: theta (device, D) ->
// out (device, D + 2), per-block partials in `partials` (blocks x (K + 3)
// doubles), on `stream`.
int plain_enqueue(sx_glm* m, const double* d_theta, double* d_out, double* partials, int propto,
                  int jacobian, hipStream_t stream, bool reduce) {
  pass_args a;
  a.X = m->d_X;
  a.y = m->d_y;
  a.theta = d_theta;
  a.partials = partials;
  a.out = d_out;
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
  a.sigma_mu = m->prior.sigma_mu;
  const double half_log_2pi = 0.91893853320467274178032973640562;
  a.alpha_const = a.alpha_sd > 0 ? -std::log(a.alpha_sd) - half_log_2pi : 0.0;
  a.beta_const = a.beta_sd > 0 ? -std::log(a.beta_sd) - half_log_2pi : 0.0;
  a.sigma_const = a.sigma_sd > 0 ? -std::log(a.sigma_sd) - half_log_2pi : 0.0;
  SX_HIP(launch_pass_f64(m->family, m->geom, a, m->blocks, m->threads, stream, reduce));
  return SX_OK;
}

int mx_chain_pad(const sx_glm* m, int C) {
  (void)m;
  if (C <= 64) return 64;
  return (C + 127) / 128 * 128;
}

int mx_prepare(sx_glm* m, mx_workspace& w, int C) {
  if (!m->d_Xh) {
    set_error("batched BF16 evaluation needs a context created with SX_BF16X2");
    return SX_EINVAL;
  }
  const int Cp = mx_chain_pad(m, C);
  const int64_t n_tiles = m->Np / 32;
  const int64_t tps = m->mx_tps > 0 ? m->mx_tps : 8;
  int splits = static_cast<int>(
      std::min<int64_t>(512, std::max<int64_t>(8, ((n_tiles + tps - 1) / tps + 7) / 8 * 8)));
  if (w.cap >= Cp && w.splits == splits) return SX_OK;
  free_workspace(w);
  const size_t ck = static_cast<size_t>(Cp) * m->Kp;
  SX_HIP(hipMalloc(&w.Bh, ck * 2));
  SX_HIP(hipMalloc(&w.Bl, ck * 2));
  SX_HIP(hipMalloc(&w.Bm, ck * 2));
  SX_HIP(hipMalloc(&w.alpha, sizeof(float) * Cp));
  SX_HIP(hipMalloc(&w.gpart, sizeof(float) * ck * splits));
  SX_HIP(hipMalloc(&w.spart, sizeof(double) * 3 * static_cast<size_t>(Cp) * splits));
  if (m->Kp > 256) {  // residuals between the forward and the adjoint launch
    const size_t nr = static_cast<size_t>(m->Np) * Cp;
    SX_HIP(hipMalloc(&w.Rh, nr * 2));
    SX_HIP(hipMalloc(&w.Rl, nr * 2));
  }
  w.cap = Cp;
  w.Cp = Cp;
  w.splits = splits;
  return SX_OK;
}

int mx_enqueue(sx_glm* m, mx_workspace& w, const double* d_theta, int C, double* d_out,
               int propto, int jacobian, hipStream_t stream) {
  int rc = mx_prepare(m, w, C);
  if (rc) return rc;
  mx_args a{};
  a.Xh = m->d_Xh;
  a.Xl = m->d_Xl;
  a.y = m->d_y;
  a.yf = m->d_yf;
  a.Bh = w.Bh;
  a.Bl = w.Bl;
  a.Bm = w.Bm;
  a.alpha = w.alpha;
  a.gpart = w.gpart;
  a.spart = w.spart;
  a.Rh = w.Rh;
  a.Rl = w.Rl;
  a.N = m->N;
  a.n_tiles = m->Np / 32;
  a.Kp = static_cast<int>(m->Kp);
  a.Cp = mx_chain_pad(m, C);
  a.splits = w.splits;
  a.tiles_per_split = (a.n_tiles + a.splits - 1) / a.splits;
  a.variant = m->mx_variant;
  pass_args pa{};
  pa.theta = d_theta;
  pa.out = d_out;
  pa.N = m->N;
  pa.K = static_cast<int>(m->K);
  pa.D = static_cast<int>(m->D);
  pa.E = static_cast<int>(m->K) + 3;
  pa.propto = propto;
  pa.jacobian = jacobian;
  pa.chains = C;
  pa.theta_stride = m->D;
  pa.out_stride = m->D + 2;
  pa.alpha_sd = m->prior.alpha_sd;
  pa.beta_sd = m->prior.beta_sd;
  pa.sigma_sd = m->prior.sigma_sd;
  pa.sigma_mu = m->prior.sigma_mu;
  const double h = 0.91893853320467274178032973640562;
  pa.alpha_const = pa.alpha_sd > 0 ? -std::log(pa.alpha_sd) - h : 0.0;
  pa.beta_const = pa.beta_sd > 0 ? -std::log(pa.beta_sd) - h : 0.0;
  pa.sigma_const = pa.sigma_sd > 0 ? -std::log(pa.sigma_sd) - h : 0.0;
  SX_HIP(launch_mfma_pass(m->family, d_theta, m->D, C, a, pa, stream));
  return SX_OK;
}

// Launch shape of the batched MFMA pass, timed on the device at context
// creation: the kernel variant (one wave per 32-chain block; a pair of
// waves split by k where Kp / 32 is even and at most 8; or, for Kp <= 128,
// one wave per block without operand prefetch at two waves per SIMD, the
// same products as the first) and the row split
// (target tiles of 32 rows per split: 4 to 128; at most 512 splits, so
// large N reach fewer splits, and fewer gradient partials to sum, only with
// the larger targets), each timed as full
// evaluations of 128 chains at theta = 0 (the kernels' work does not depend
// on the values).  The split sets the order of the FP64 sums, so, as for
// the FP64 pass, the first decision per (device, family, N, K) is kept for
// the process (and in STANX_TUNING_CACHE, as family + 100).  The choice
// does not depend on the batch size, so a chain's result stays independent
// of the other chains in its batch.
int mx_autotune(sx_glm* m) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  read_tune_file();
  const std::string key = tune_key(m->device, 100 + m->family, m->N, m->K);
  const auto hit = g_tuned.find(key);
  if (hit != g_tuned.end()) {
    m->mx_variant = hit->second.L;
    m->mx_tps = hit->second.smax;
    char line[160];
    std::snprintf(line, sizeof line, "MFMA pass cached: variant=%d tiles/split=%d\n",
                  m->mx_variant, m->mx_tps);
    m->tuning_log += line;
    return SX_OK;
  }
  const int C = 128;
  const int64_t D = m->D;
  double *d_th = nullptr, *d_out = nullptr;
  if (hipMalloc(&d_th, sizeof(double) * C * D) != hipSuccess ||
      hipMalloc(&d_out, sizeof(double) * C * (D + 2)) != hipSuccess) {
    (void)hipFree(d_th);
    set_error("MFMA autotune: out of device memory");
    return SX_ENOMEM;
  }
  mx_workspace w;
  std::string log;
  int rc = hipMemsetAsync(d_th, 0, sizeof(double) * C * D, m->stream) == hipSuccess
               ? SX_OK : SX_EDEVICE;
  const int kq = static_cast<int>(m->Kp / 32);
  int best_v = 0, best_t = 8;
  double best_ms = 1e30;
  for (int v = 0; v <= 2 && rc == SX_OK; ++v) {
    if (v == 1 && (m->Kp > 256 || kq % 2 != 0)) continue;
    if (v == 2 && kq > 4) continue;
    for (int tps : {4, 8, 16, 32, 64, 128}) {
      m->mx_variant = v;
      m->mx_tps = tps;
      for (int i = 0; i < 2 && rc == SX_OK; ++i)
        rc = mx_enqueue(m, w, d_th, C, d_out, 1, 1, m->stream);
      if (rc) break;
      if (hipEventRecord(m->ev0, m->stream) != hipSuccess) { rc = SX_EDEVICE; break; }
      const int reps = 3;
      for (int i = 0; i < reps && rc == SX_OK; ++i)
        rc = mx_enqueue(m, w, d_th, C, d_out, 1, 1, m->stream);
      if (rc) break;
      float ms = 0.f;
      if (hipEventRecord(m->ev1, m->stream) != hipSuccess ||
          hipEventSynchronize(m->ev1) != hipSuccess ||
          hipEventElapsedTime(&ms, m->ev0, m->ev1) != hipSuccess) {
        rc = SX_EDEVICE;
        break;
      }
      const double per = ms / reps;
      char line[160];
      std::snprintf(line, sizeof line, "MFMA pass variant=%d tiles/split=%d: %.4f ms (128 chains)\n",
                    v, tps, per);
      log += line;
      if (per < best_ms) {
        best_ms = per;
        best_v = v;
        best_t = tps;
      }
    }
  }
  (void)hipStreamSynchronize(m->stream);
  free_workspace(w);
  (void)hipFree(d_th);
  (void)hipFree(d_out);
  m->mx_variant = best_v;
  m->mx_tps = best_t;
  if (rc) return rc;
  remember_tuning(key, tuned{best_v, best_t, 0, 0});
  char line[160];
  std::snprintf(line, sizeof line, "MFMA pass chosen: variant=%d tiles/split=%d (%.4f ms)\n",
                best_v, best_t, best_ms);
  log += line;
  m->tuning_log += log;
  return SX_OK;
}

// Hierarchical model, C chains: theta (device, C x D) -> out (device,
// C x (D + 2)); FP64 X, or FP32 X with chains on lanes.
int hier_prepare(sx_glm* m, hier_workspace& w, int C) {
  if (w.cap >= C) return SX_OK;
  free_workspace(w);
  const size_t E = static_cast<size_t>(m->K) + 3, G = 2 * static_cast<size_t>(m->K) + 5;
  SX_HIP(hipMalloc(&w.cpart, sizeof(double) * E * std::max<int64_t>(1, m->n_chunks) * C));
  SX_HIP(hipMalloc(&w.gpart, sizeof(double) * G * m->J * C));
  if (m->precision == SX_FP32 && m->K <= 31) {
    // the coefficient kernel writes columns 0 .. K; the rest stay zero
    SX_HIP(hipMalloc(&w.wsplit, sizeof(uint16_t) * 96 * m->J * C));
    SX_HIP(hipMemset(w.wsplit, 0, sizeof(uint16_t) * 96 * m->J * C));
  }
  w.cap = C;
  return SX_OK;
}

int hier_enqueue(sx_glm* m, hier_workspace& w, const double* d_theta, int C, double* d_out,
                 int propto, int jacobian, bool fp32, hipStream_t stream) {
  if (fp32 && !m->d_X32 && m->N * m->K > 0) {
    set_error("FP32 evaluation needs a context created with SX_FP32");
    return SX_EINVAL;
  }
  int rc = hier_prepare(m, w, C);
  if (rc) return rc;
  hier_args a{};
  a.X = m->d_X;
  a.X32 = m->d_X32;
  a.XS = m->d_XS;
  a.y = static_cast<const int8_t*>(m->d_y);
  a.theta = d_theta;
  a.out = d_out;
  a.cpart = w.cpart;
  a.gpart = w.gpart;
  a.wsplit = w.wsplit;
  a.chunk_group = m->d_chunk_group;
  a.chunk_row0 = m->d_chunk_row0;
  a.group_chunk0 = m->d_group_chunk0;
  a.n_chunks = m->n_chunks;
  a.theta_stride = m->D;
  a.out_stride = m->D + 2;
  a.J = static_cast<int>(m->J);
  a.K = static_cast<int>(m->K);
  a.D = static_cast<int>(m->D);
  a.E = static_cast<int>(m->K) + 3;
  a.G = 2 * static_cast<int>(m->K) + 5;
  a.chains = C;
  a.propto = propto;
  a.jacobian = jacobian;
  a.x6 = m->hier_x6;
  a.alpha_sd = m->prior.alpha_sd;
  a.beta_sd = m->prior.beta_sd;
  a.tau_sd = m->prior.tau_sd;
  const double h = 0.91893853320467274178032973640562;
  a.alpha_const = a.alpha_sd > 0 ? -std::log(a.alpha_sd) - h : 0.0;
  a.beta_const = a.beta_sd > 0 ? -std::log(a.beta_sd) - h : 0.0;
  a.tau_const = a.tau_sd > 0 ? -std::log(a.tau_sd) - h : 0.0;
  SX_HIP(launch_hier(a, fp32, stream));
  return SX_OK;
}

}  // namespace sx

namespace {

int ensure_batch_staging(sx_glm* m, int64_t C) {
  if (m->bcap >= C) return SX_OK;
  (void)hipFree(m->d_btheta);
  (void)hipFree(m->d_bout);
  m->d_btheta = nullptr;
  m->d_bout = nullptr;
  m->bcap = 0;
  SX_HIP(hipMalloc(&m->d_btheta, sizeof(double) * C * m->D));
  SX_HIP(hipMalloc(&m->d_bout, sizeof(double) * C * (m->D + 2)));
  m->bcap = C;
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
  if (C == 0) return SX_OK;
  if (m->precision == SX_BF16X2 || m->family == SX_HIER_LOGIT) {
    SX_HIP(hipSetDevice(m->device));
    int rc = ensure_batch_staging(m, C);
    if (rc) return rc;
    const int64_t D = m->D;
    SX_HIP(hipMemcpyAsync(m->d_btheta, theta, sizeof(double) * C * D, hipMemcpyHostToDevice,
                          m->stream));
    rc = m->family == SX_HIER_LOGIT
             ? sx::hier_enqueue(m, m->hw, m->d_btheta, static_cast<int>(C), m->d_bout, propto,
                                jacobian, m->precision == SX_FP32, m->stream)
             : sx::mx_enqueue(m, m->mx, m->d_btheta, static_cast<int>(C), m->d_bout, propto,
                              jacobian, m->stream);
    if (rc) return rc;
    std::vector<double> out(static_cast<size_t>(C * (D + 2)));
    SX_HIP(hipMemcpyAsync(out.data(), m->d_bout, sizeof(double) * out.size(),
                          hipMemcpyDeviceToHost, m->stream));
    SX_HIP(hipStreamSynchronize(m->stream));
    int worst = SX_OK;
    for (int64_t c = 0; c < C; ++c) {
      const double* o = out.data() + c * (D + 2);
      lp[c] = o[0];
      if (grad) std::memcpy(grad + c * D, o + 1, sizeof(double) * D);
      if (o[D + 1] != 0.0) {
        lp[c] = -INFINITY;
        worst = SX_EDOMAIN;
        set_error(m->family == SX_HIER_LOGIT
                      ? "bernoulli_logit_lpmf: linear predictor or log density is not finite"
                      : "batched GLM: linear predictor or log density is not finite");
      }
    }
    return worst;
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
  if (!m || !theta || iters <= 0 || C < 1 ||
      (C > 1 && m->precision != SX_BF16X2 && m->family != SX_HIER_LOGIT)) {
    set_error("sx_glm_time_logp_grad: invalid argument");
    return SX_EINVAL;
  }
  SX_HIP(hipSetDevice(m->device));
  const size_t D = static_cast<size_t>(m->D);
  const bool batched = C > 1;
  if (batched) {
    int rc = ensure_batch_staging(m, C);
    if (rc) return rc;
    SX_HIP(hipMemcpy(m->d_btheta, theta, D * C * sizeof(double), hipMemcpyHostToDevice));
  } else {
    SX_HIP(hipMemcpy(m->d_theta, theta, D * sizeof(double), hipMemcpyHostToDevice));
  }
  auto one = [&](bool reduce) -> int {
    if (batched && m->family == SX_HIER_LOGIT)
      return sx::hier_enqueue(m, m->hw, m->d_btheta, static_cast<int>(C), m->d_bout, 1, 1,
                              m->precision == SX_FP32, m->stream);
    if (batched)
      return sx::mx_enqueue(m, m->mx, m->d_btheta, static_cast<int>(C), m->d_bout, 1, 1, m->stream);
    return enqueue_pass(m, 1, 1, reduce);
  };
  for (int i = 0; i < warmup; ++i) {
    int rc = one(true);
    if (rc) return rc;
  }
  SX_HIP(hipStreamSynchronize(m->stream));
  // full evaluations (for a single chain: bulk pass + cross-block reduction)
  SX_HIP(hipEventRecord(m->ev0, m->stream));
  for (int i = 0; i < iters; ++i) {
    int rc = one(true);
    if (rc) return rc;
  }
  SX_HIP(hipEventRecord(m->ev1, m->stream));
  SX_HIP(hipEventSynchronize(m->ev1));
  float ms = 0.f;
  SX_HIP(hipEventElapsedTime(&ms, m->ev0, m->ev1));
  if (ms_per_eval) *ms_per_eval = ms / iters;
  if (batched) {
    if (ms_kernel) *ms_kernel = ms / iters;
    return SX_OK;
  }
  // the bulk X pass alone (its partials are left unreduced)
  SX_HIP(hipEventRecord(m->ev0, m->stream));
  for (int i = 0; i < iters; ++i) {
    int rc = one(false);
    if (rc) return rc;
  }
  SX_HIP(hipEventRecord(m->ev1, m->stream));
  SX_HIP(hipEventSynchronize(m->ev1));
  SX_HIP(hipEventElapsedTime(&ms, m->ev0, m->ev1));
  if (ms_kernel) *ms_kernel = ms / iters;
  return SX_OK;
}

int sx_glm_kernel_config(const sx_glm* m, int* lanes_per_row, int* slots_per_lane,
                         int* blocks, int* threads, int* prefetch) {
  if (!m) {
    set_error("NULL model");
    return SX_EINVAL;
  }
  if (lanes_per_row) *lanes_per_row = m->geom.L;
  if (slots_per_lane) *slots_per_lane = m->geom.S;
  if (blocks) *blocks = m->blocks;
  if (threads) *threads = m->threads;
  if (prefetch) *prefetch = m->geom.prefetch;
  return SX_OK;
}

}  // extern "C"
