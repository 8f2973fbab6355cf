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
// Block partials are summed in a fixed order by a second one-block launch,
// so results are bitwise reproducible, and that launch adds the priors and
// the log-Jacobian and writes lp and the gradient.  This replaces the two GEMV passes over X that the reference's
// CPU GLM performs per gradient (Stan Math bernoulli_logit_glm_lpmf /
// normal_id_glm_lpdf, reached through src/stan/model/gradient.hpp:21-35).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../include/stanx/sx_math.hpp"
#include "glm_finalize.hpp"
#include "sx_internal.hpp"

namespace sx {

namespace {

// Buffer-resource loads: lanes whose offset is past `bytes` read zero and
// generate no memory request, so a tile's SMAX slot loads need no
// branches when the layout only holds S <= SMAX slots.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base,
                                                            uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), static_cast<short>(0),
                                           static_cast<int>(bytes), 0x00020000);
}
__device__ __forceinline__ double buf_f64(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                          uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 2));
}

template <int SMAX>
__device__ __forceinline__ void load_tile(const pass_args& a, int64_t t, int lane,
                                          double (&x)[SMAX]) {
  const __amdgpu_buffer_rsrc_t rs =
      make_rsrc(a.X + t * static_cast<int64_t>(a.S) * 64, static_cast<uint32_t>(a.S) * 512u);
#pragma unroll
  for (int j = 0; j < SMAX; ++j) x[j] = buf_f64(rs, lane * 8u, j * 512u);
}

// y of this lane's row (zero past N); issued before the tile's X loads so
// that waiting for it leaves the X loads in flight
template <int FAM>
__device__ __forceinline__ double load_y(const pass_args& a, int64_t t, int R, int r) {
  const int64_t row0 = t * R;
  const int64_t rows = a.N - row0 < R ? a.N - row0 : R;
  if (FAM == SX_LOGIT) {
    const __amdgpu_buffer_rsrc_t rs =
        make_rsrc(static_cast<const int8_t*>(a.y) + row0, static_cast<uint32_t>(rows));
    return static_cast<double>(__builtin_amdgcn_raw_buffer_load_b8(rs, r, 0, 0));
  } else {
    const __amdgpu_buffer_rsrc_t rs =
        make_rsrc(static_cast<const double*>(a.y) + row0, static_cast<uint32_t>(rows) * 8u);
    return buf_f64(rs, r * 8u, 0);
  }
}

template <int SMAX, int FAM, int PF>
__global__ __launch_bounds__(256) void glm_pass_f64(pass_args a) {
  extern __shared__ double lds[];
  // 1-D grid of chains x nb blocks.  With nb a multiple of 8 the blocks of
  // every chain that stream the same tiles sit on one XCD (block ids 8
  // apart), so a batched evaluation reads each tile of X from HBM about
  // once per XCD and the other chains find it in that XCD's L2; for one
  // chain the mapping is the identity.
  const int nb = static_cast<int>(gridDim.x) / a.chains;
  int blk, chain;
  {
    const int L = blockIdx.x;
    if ((nb & 7) == 0) {
      const int q = L >> 3;
      blk = (L & 7) + 8 * (q / a.chains);
      chain = q % a.chains;
    } else {
      blk = L % nb;
      chain = L / nb;
    }
    blk = __builtin_amdgcn_readfirstlane(blk);
    chain = __builtin_amdgcn_readfirstlane(chain);
  }
  a.theta += static_cast<int64_t>(chain) * a.theta_stride;
  a.partials += static_cast<int64_t>(chain) * a.E * nb;
  const int lane = threadIdx.x & 63;
  // the wave index is uniform across the wave: say so, so that tile
  // indices and buffer descriptors live in scalar registers
  const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = blockDim.x >> 6;
  const int L = a.L;
  const int q = lane & (L - 1);
  const int r = lane >> a.lgL;
  const int R = 64 >> a.lgL;
  const int S = a.S;

  // beta staged in LDS, padded to SMAX*L entries (zero beyond K)
  double* sbeta = lds + nw * a.E;
  for (int i = threadIdx.x; i < SMAX * L; i += blockDim.x)
    sbeta[i] = i < a.K ? a.theta[1 + i] : 0.0;
  __syncthreads();
  const double alpha = a.theta[0];
  double acc[SMAX];
#pragma unroll
  for (int j = 0; j < SMAX; ++j) acc[j] = 0.0;
  double s0 = 0.0, s1 = 0.0, bad = 0.0;

  const int64_t W = static_cast<int64_t>(nb) * nw;
  int64_t t = static_cast<int64_t>(blk) * nw + wib;
  double xv[SMAX];
  double yv = 0.0;
  if (PF && t < a.T) {
    yv = load_y<FAM>(a, t, R, r);
    load_tile<SMAX>(a, t, lane, xv);
  }
  for (; t < a.T; t += W) {
    double xn[SMAX];
    double yn = 0.0;
    if (PF) {
      // the next tile's loads are in flight during this tile's arithmetic
      const int64_t tn = t + W;
      if (tn < a.T) {
        yn = load_y<FAM>(a, tn, R, r);
        load_tile<SMAX>(a, tn, lane, xn);
      } else {
#pragma unroll
        for (int j = 0; j < SMAX; ++j) xn[j] = 0.0;
      }
    } else {
      yv = load_y<FAM>(a, t, R, r);
      load_tile<SMAX>(a, t, lane, xv);
    }
    // beta is re-read from LDS every tile: an opaque zero offset keeps the
    // compiler from hoisting SMAX loop-invariant values into registers
    int zoff = 0;
    asm volatile("" : "+s"(zoff));
    const double* sb = sbeta + zoff + q;
    double e = 0.0;
#pragma unroll
    for (int j = 0; j < SMAX; ++j) e = fma(xv[j], sb[j * L], e);
    for (int off = 1; off < L; off <<= 1) e += __shfl_xor(e, off, 64);
    const int64_t row = t * R + r;
    const bool valid = row < a.N;
    e += alpha;
    double rr = 0.0, term = 0.0;
    if (valid) {
      if (FAM == SX_LOGIT) {
        stanx::math::sx_logit_row(e, yv != 0.0 ? 1.0 : -1.0, &term, &rr);
      } else {
        const double d = yv - e;
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
    if (PF) {
#pragma unroll
      for (int j = 0; j < SMAX; ++j) xv[j] = xn[j];
      yv = yn;
    }
  }

  // reduce the per-lane column partials over the R row-lanes of each column
  for (int off = L; off < 64; off <<= 1) {
#pragma unroll
    for (int j = 0; j < SMAX; ++j)
      if (j < S) acc[j] += __shfl_xor(acc[j], off, 64);
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
  // block partial, stored entry-major: partials[i * nb + block]
  for (int i = threadIdx.x; i < E; i += blockDim.x) {
    double v = 0.0;
    for (int w = 0; w < nw; ++w) v += lds[w * E + i];
    a.partials[static_cast<int64_t>(i) * nb + blk] = v;
  }
}

// Cross-block reduction of the bulk pass's partials (fixed order, so the
// result is bitwise reproducible for a given launch shape) followed by the
// priors, Jacobian and output.  A separate launch: the kernel boundary
// makes every block's partials visible without per-block cache
// write-backs.  One block of 1024 threads; wave w sums entries w, w+16, ...
template <int FAM>
__global__ __launch_bounds__(1024) void glm_reduce_f64(pass_args a, int nb) {
  extern __shared__ double st[];
  a.theta += static_cast<int64_t>(blockIdx.x) * a.theta_stride;
  a.partials += static_cast<int64_t>(blockIdx.x) * a.E * nb;
  a.out += static_cast<int64_t>(blockIdx.x) * a.out_stride;
  const int lane = threadIdx.x & 63;
  const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nw = blockDim.x >> 6;
  const int E = a.E;
  // Each wave takes 8 entries per round; every lane issues all of its
  // loads (8 entries x 4 blocks per step) before any add: clamped indices
  // and 0/1 weights keep the loop free of branches.
  for (int i0 = wib * 8; i0 < E; i0 += nw * 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = 0.0;
    for (int b0 = 0; b0 < nb; b0 += 256) {
      double x[8][4];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = min(i0 + u, E - 1);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int b = min(b0 + lane + 64 * k, nb - 1);
          x[u][k] = a.partials[static_cast<int64_t>(i) * nb + b];
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int k = 0; k < 4; ++k)
          v[u] += (b0 + lane + 64 * k < nb) ? x[u][k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const double w = wave_sum(v[u]);
      if (lane == 0 && i0 + u < E) st[i0 + u] = w;
    }
  }
  __syncthreads();
  if (wib == 0) finalize<FAM>(a, st);
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

template <int SMAX, int FAM, int PF>
hipError_t launch_one(const pass_args& a, int blocks, int threads, hipStream_t s,
                      bool reduce) {
  const size_t lds =
      (static_cast<size_t>(threads / 64) * a.E + static_cast<size_t>(SMAX) * a.L) * sizeof(double);
  hipLaunchKernelGGL((glm_pass_f64<SMAX, FAM, PF>), dim3(blocks * a.chains), dim3(threads),
                     lds, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !reduce) return e;
  hipLaunchKernelGGL((glm_reduce_f64<FAM>), dim3(a.chains), dim3(1024), a.E * sizeof(double),
                     s, a, blocks);
  return hipGetLastError();
}

template <int FAM, int PF>
hipError_t launch_fam(int smax, const pass_args& a, int blocks, int threads, hipStream_t s,
                      bool reduce) {
  switch (smax) {
    case 4: return launch_one<4, FAM, PF>(a, blocks, threads, s, reduce);
    case 8: return launch_one<8, FAM, PF>(a, blocks, threads, s, reduce);
    case 12: return launch_one<12, FAM, PF>(a, blocks, threads, s, reduce);
    case 16: return launch_one<16, FAM, PF>(a, blocks, threads, s, reduce);
    case 20: return launch_one<20, FAM, PF>(a, blocks, threads, s, reduce);
    case 24: return launch_one<24, FAM, PF>(a, blocks, threads, s, reduce);
    case 28: return launch_one<28, FAM, PF>(a, blocks, threads, s, reduce);
    case 32: return launch_one<32, FAM, PF>(a, blocks, threads, s, reduce);
  }
  return hipErrorInvalidValue;
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
                           int blocks, int threads, hipStream_t s, bool reduce) {
  if (family == SX_LOGIT)
    return g.prefetch ? launch_fam<SX_LOGIT, 1>(g.smax, a, blocks, threads, s, reduce)
                      : launch_fam<SX_LOGIT, 0>(g.smax, a, blocks, threads, s, reduce);
  if (family == SX_NORMAL)
    return g.prefetch ? launch_fam<SX_NORMAL, 1>(g.smax, a, blocks, threads, s, reduce)
                      : launch_fam<SX_NORMAL, 0>(g.smax, a, blocks, threads, s, reduce);
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

}  // namespace sx
