// Fused FP64 forward + adjoint pass of the GLM log density on gfx950.
//
// One launch evaluates, for theta = (alpha, beta[, u]):
//   logistic  (bernoulli_logit_glm):  eta = X beta + alpha,
//             lp_lik = sum_n -log1p(exp(-s_n eta_n)),  s_n = 2 y_n - 1,
//             r_n = s_n / (1 + exp(s_n eta_n)) = y_n - inv_logit(eta_n)
//   normal    (normal_id_glm):        d_n = y_n - (X beta + alpha)_n,
//             SS = sum d_n^2, r_n = d_n
// and the adjoint X^T r, sum r, reading every element of X exactly once:
// each wave loads a 64-row-block tile of X into registers (one 512-byte
// coalesced load per slot), forms eta with an L-lane shuffle reduction,
// evaluates the link, and accumulates x * r into per-lane FP64 registers.
// Block partials are summed in a fixed order by the last block to finish
// (ticket counter), so results are bitwise reproducible, and the same
// block adds the priors and the log-Jacobian and writes lp and the
// gradient.  This replaces the two GEMV passes over X that the reference's
// CPU GLM performs per gradient (Stan Math bernoulli_logit_glm_lpmf /
// normal_id_glm_lpdf, reached through src/stan/model/gradient.hpp:21-35).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "sx_internal.hpp"

namespace sx {

namespace {

constexpr double kHalfLog2Pi = 0.91893853320467274178032973640562;

__device__ __forceinline__ double log1p_exp_d(double x) {
  return x > 0.0 ? x + log1p(exp(-x)) : log1p(exp(x));
}

__device__ __forceinline__ double load_coherent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of v over the 64 lanes (xor butterfly): every lane gets the total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Adds priors, Jacobian and scaling to the reduced statistics st[0..E) and
// writes out[0] = lp, out[1..D] = gradient, out[D+1] = status (0 ok,
// 1 domain error).  Runs on the first wave of the last block.
template <int FAM>
__device__ void finalize(const pass_args& a, const double* st) {
  const int lane = threadIdx.x & 63;
  const int K = a.K;
  const double* th = a.theta;
  const double alpha = th[0];
  double lp_prior = 0.0;  // lane-partial of prior terms over beta
  double scale = 1.0;
  if (FAM == SX_NORMAL) {
    const double sigma = exp(th[K + 1]);
    scale = 1.0 / (sigma * sigma);
  }
  for (int k = lane; k < K; k += 64) {
    const double b = th[1 + k];
    double g = st[2 + k] * scale;
    if (a.beta_sd > 0) {
      const double z = b / a.beta_sd;
      lp_prior += -0.5 * z * z;
      if (!a.propto) lp_prior += -log(a.beta_sd) - kHalfLog2Pi;
      g += -b / (a.beta_sd * a.beta_sd);
    }
    a.out[2 + k] = g;
  }
  lp_prior = wave_sum(lp_prior);
  if (lane != 0) return;
  double lp = 0.0;
  double ga = st[1] * scale;
  double bad = st[K + 2];
  if (FAM == SX_LOGIT) {
    lp = st[0];
  } else {
    const double u = th[K + 1];
    const double sigma = exp(u);
    if (!(sigma > 0) || !isfinite(sigma)) bad += 1.0;
    const double SS = st[0];
    lp = -0.5 * SS * scale - static_cast<double>(a.N) * u;
    if (!a.propto) lp += -static_cast<double>(a.N) * kHalfLog2Pi;
    double gu = -static_cast<double>(a.N) + SS * scale;
    if (a.sigma_sd > 0) {
      const double z = sigma / a.sigma_sd;
      lp += -0.5 * z * z;
      if (!a.propto) lp += -log(a.sigma_sd) - kHalfLog2Pi;
      gu += -z * z;
    }
    if (a.jacobian) {
      lp += u;
      gu += 1.0;
    }
    a.out[K + 2] = gu;
  }
  if (a.alpha_sd > 0) {
    const double z = alpha / a.alpha_sd;
    lp += -0.5 * z * z;
    if (!a.propto) lp += -log(a.alpha_sd) - kHalfLog2Pi;
    ga += -alpha / (a.alpha_sd * a.alpha_sd);
  }
  lp += lp_prior;
  a.out[0] = lp;
  a.out[1] = ga;
  if (!isfinite(lp)) bad += 1.0;
  a.out[a.D + 1] = bad > 0 ? 1.0 : 0.0;
}

template <int SMAX, int FAM>
__global__ __launch_bounds__(256) void glm_pass_f64(pass_args a) {
  extern __shared__ double lds[];
  __shared__ int is_last;
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  const int L = a.L;
  const int q = lane & (L - 1);
  const int r = lane >> a.lgL;
  const int R = 64 >> a.lgL;
  const int S = a.S;

  // beta staged in LDS, padded to S*L entries (zero beyond K)
  double* sbeta = lds + nw * a.E;
  for (int i = threadIdx.x; i < S * L; i += blockDim.x)
    sbeta[i] = i < a.K ? a.theta[1 + i] : 0.0;
  __syncthreads();
  const double alpha = a.theta[0];
  double acc[SMAX];
#pragma unroll
  for (int j = 0; j < SMAX; ++j) acc[j] = 0.0;
  double s0 = 0.0, s1 = 0.0, bad = 0.0;

  const int64_t W = static_cast<int64_t>(gridDim.x) * nw;
  const int64_t tile_stride = static_cast<int64_t>(S) * 64;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * nw + wib; t < a.T; t += W) {
    const double* base = a.X + t * tile_stride + lane;
    double xv[SMAX];
#pragma unroll
    for (int j = 0; j < SMAX; ++j)
      xv[j] = (j < S) ? __builtin_nontemporal_load(base + j * 64) : 0.0;
    double e = 0.0;
#pragma unroll
    for (int j = 0; j < SMAX; ++j)
      if (j < S) e = fma(xv[j], sbeta[j * L + q], e);
    for (int off = 1; off < L; off <<= 1) e += __shfl_xor(e, off, 64);
    const int64_t row = t * R + r;
    const bool valid = row < a.N;
    e += alpha;
    double rr = 0.0, term = 0.0;
    if (valid) {
      if (FAM == SX_LOGIT) {
        const int yv = static_cast<const int8_t*>(a.y)[row];
        const double s = yv ? 1.0 : -1.0;
        const double se = s * e;
        rr = s / (1.0 + exp(se));
        term = -log1p_exp_d(-se);
      } else {
        const double d = static_cast<const double*>(a.y)[row] - e;
        rr = d;
        term = d * d;
      }
      if (!isfinite(e)) bad += 1.0;
    }
    if (q == 0) {
      s0 += term;
      s1 += rr;
    }
#pragma unroll
    for (int j = 0; j < SMAX; ++j) acc[j] = fma(xv[j], rr, acc[j]);
  }

  // reduce the per-lane column partials over the R row-lanes of each column
  for (int off = L; off < 64; off <<= 1) {
#pragma unroll
    for (int j = 0; j < SMAX; ++j) acc[j] += __shfl_xor(acc[j], off, 64);
  }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  bad = wave_sum(bad);
  const int E = a.E;
  if (r == 0) {
#pragma unroll
    for (int j = 0; j < SMAX; ++j) {
      const int c = j * L + q;
      if (j < S && c < a.K) lds[wib * E + 2 + c] = acc[j];
    }
  }
  if (lane == 0) {
    lds[wib * E + 0] = s0;
    lds[wib * E + 1] = s1;
    lds[wib * E + a.K + 2] = bad;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < E; i += blockDim.x) {
    double v = 0.0;
    for (int w = 0; w < nw; ++w) v += lds[w * E + i];
    a.partials[static_cast<int64_t>(blockIdx.x) * E + i] = v;
  }
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned ticket = atomicAdd(a.counter, 1u);
    is_last = (ticket == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return;
  __threadfence();
  // fixed-order sum of the block partials; four interleaved chains keep
  // several loads in flight per thread
  for (int i = threadIdx.x; i < E; i += blockDim.x) {
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
    const int nb = gridDim.x;
    int b = 0;
    for (; b + 3 < nb; b += 4) {
      v0 += load_coherent(a.partials + static_cast<int64_t>(b) * E + i);
      v1 += load_coherent(a.partials + static_cast<int64_t>(b + 1) * E + i);
      v2 += load_coherent(a.partials + static_cast<int64_t>(b + 2) * E + i);
      v3 += load_coherent(a.partials + static_cast<int64_t>(b + 3) * E + i);
    }
    for (; b < nb; ++b) v0 += load_coherent(a.partials + static_cast<int64_t>(b) * E + i);
    lds[i] = (v0 + v1) + (v2 + v3);
  }
  __syncthreads();
  if (threadIdx.x == 0) *a.counter = 0u;
  if (wib == 0) finalize<FAM>(a, lds);
}

__global__ void tile_x_f64(const double* __restrict__ Xrm, double* __restrict__ Xt,
                           int64_t N, int64_t K, int L, int lgL, int S, int64_t total) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int lane = static_cast<int>(idx & 63);
  const int64_t js = idx >> 6;  // t * S + j
  const int64_t t = js / S;
  const int j = static_cast<int>(js - t * S);
  const int64_t row = t * (64 >> lgL) + (lane >> lgL);
  const int64_t col = static_cast<int64_t>(j) * L + (lane & (L - 1));
  Xt[idx] = (row < N && col < K) ? Xrm[row * K + col] : 0.0;
}

template <int SMAX, int FAM>
hipError_t launch_one(const pass_args& a, int blocks, int threads, hipStream_t s) {
  const size_t lds =
      (static_cast<size_t>(threads / 64) * a.E + static_cast<size_t>(a.S) * a.L) * sizeof(double);
  hipLaunchKernelGGL((glm_pass_f64<SMAX, FAM>), dim3(blocks), dim3(threads), lds, s, a);
  return hipGetLastError();
}

template <int FAM>
hipError_t launch_fam(int smax, const pass_args& a, int blocks, int threads,
                      hipStream_t s) {
  switch (smax) {
    case 4: return launch_one<4, FAM>(a, blocks, threads, s);
    case 8: return launch_one<8, FAM>(a, blocks, threads, s);
    case 16: return launch_one<16, FAM>(a, blocks, threads, s);
    case 32: return launch_one<32, FAM>(a, blocks, threads, s);
  }
  return hipErrorInvalidValue;
}

template <int SMAX, int FAM>
int occ_one(int threads) {
  int nb = 0;
  const size_t lds = (static_cast<size_t>(threads / 64) * 515 + 1024) * sizeof(double);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, glm_pass_f64<SMAX, FAM>, threads,
                                                   lds) != hipSuccess)
    return 1;
  return nb > 0 ? nb : 1;
}

}  // namespace

tile_geom make_geom(int64_t N, int64_t K, int L, int smax) {
  tile_geom g;
  g.L = L;
  g.lgL = 0;
  while ((1 << g.lgL) < L) ++g.lgL;
  g.R = 64 / L;
  g.S = static_cast<int>((K + L - 1) / L);
  g.smax = smax;
  g.T = (N + g.R - 1) / g.R;
  return g;
}

hipError_t launch_pass_f64(int family, const tile_geom& g, const pass_args& a,
                           int blocks, int threads, hipStream_t s) {
  if (family == SX_LOGIT) return launch_fam<SX_LOGIT>(g.smax, a, blocks, threads, s);
  if (family == SX_NORMAL) return launch_fam<SX_NORMAL>(g.smax, a, blocks, threads, s);
  return hipErrorInvalidValue;
}

hipError_t launch_tile_x_f64(const double* Xrm, double* Xt, int64_t N, int64_t K,
                             const tile_geom& g, hipStream_t s) {
  const int64_t total = g.T * g.S * 64;
  if (total == 0) return hipSuccess;
  const int threads = 256;
  const int64_t blocks = (total + threads - 1) / threads;
  hipLaunchKernelGGL(tile_x_f64, dim3(static_cast<unsigned>(blocks)), dim3(threads), 0, s,
                     Xrm, Xt, N, K, g.L, g.lgL, g.S, total);
  return hipGetLastError();
}

int max_blocks_per_cu(int family, int smax, int threads) {
  if (family == SX_NORMAL) {
    switch (smax) {
      case 4: return occ_one<4, SX_NORMAL>(threads);
      case 8: return occ_one<8, SX_NORMAL>(threads);
      case 16: return occ_one<16, SX_NORMAL>(threads);
      default: return occ_one<32, SX_NORMAL>(threads);
    }
  }
  switch (smax) {
    case 4: return occ_one<4, SX_LOGIT>(threads);
    case 8: return occ_one<8, SX_LOGIT>(threads);
    case 16: return occ_one<16, SX_LOGIT>(threads);
    default: return occ_one<32, SX_LOGIT>(threads);
  }
}

}  // namespace sx
