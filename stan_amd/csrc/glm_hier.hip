// Hierarchical logistic regression with ragged groups (segmented
// reductions), single chain in FP64 and many chains in FP32.
//
// Model (theta = [alpha, beta_1..K, v_0..v_K, z_{1,0..K} .. z_{J,0..K}],
// tau_k = exp(v_k)):
//   eta_n = alpha + tau_0 z_{g,0} + sum_k x_nk (beta_k + tau_k z_{g,k}),  g = group(n)
//   y_n ~ bernoulli_logit(eta_n), z ~ normal(0, 1), tau ~ half-normal(0, tau_sd),
//   alpha ~ normal(0, alpha_sd), beta ~ normal(0, beta_sd), log-Jacobian sum v_k.
// The gradient needs the segment sums S_{g,0} = sum_{n in g} r_n and
// S_{g,k} = sum_{n in g} x_nk r_n (r = y - inv_logit(eta)):
//   d/dz_{g,k} = tau_k S_{g,k} - z_{g,k},  d/dbeta_k = sum_g S_{g,k},
//   d/dalpha = sum_g S_{g,0},  d/dv_k = tau_k sum_g z_{g,k} S_{g,k} - tau_k^2/tau_sd^2 + 1.
// (Specification analogue: src/test/test-models/good/variational/hier_logistic.stan.)
//
// Rows are cut into chunks of at most `chunk` rows that never straddle a
// group, so ragged group sizes become uniform work items and the largest
// group does not serialise the pass.  Every chunk yields FP64 partials
// [S_0..S_K, lp, bad]; a per-group launch sums its chunks in order (z
// gradients), a final launch sums over groups (alpha, beta, v, priors).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../include/stanx/sx_math.hpp"
#include "glm_finalize.hpp"
#include "sx_internal.hpp"

namespace sx {

namespace {

constexpr double kHalfLog2PiH = 0.91893853320467274178032973640562;

// FP64: a wave per chunk, rows on lanes; chains on grid.y.  A chunk's sums
// are formed by one wave in a fixed order, so each chain's result does not
// depend on the launch shape or on the other chains.
template <int KMAX>
__global__ __launch_bounds__(256) void hier_pass_f64(hier_args a) {
  __shared__ double sw[4][KMAX + 1];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int K = a.K;
  const int c = blockIdx.y;
  const double* th = a.theta + static_cast<int64_t>(c) * a.theta_stride;
  const int64_t W = static_cast<int64_t>(gridDim.x) * 4;
  for (int64_t ch = static_cast<int64_t>(blockIdx.x) * 4 + w; ch < a.n_chunks; ch += W) {
    const int g = a.chunk_group[ch];
    const int64_t r0 = a.chunk_row0[ch], r1 = a.chunk_row0[ch + 1];
    const double* z = th + 2 + 2 * K + static_cast<int64_t>(g) * (K + 1);
    if (lane <= K) {
      const double tau = exp(th[1 + K + lane]);
      sw[w][lane] = lane == 0 ? th[0] + tau * z[0] : th[lane] + tau * z[lane];
    }
    // sw[w] is this wave's: wait for the LDS writes, then a wave barrier
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    double acc[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) acc[k] = 0.0;
    double s0 = 0.0, s1 = 0.0, bad = 0.0;
    for (int64_t base = r0; base < r1; base += 64) {
      const int64_t n = base + lane;
      const bool valid = n < r1;
      const double* x = a.X + n * K;
      double e = sw[w][0];
      double xv[KMAX];
#pragma unroll
      for (int k = 0; k < KMAX; ++k) {
        xv[k] = (valid && k < K) ? x[k] : 0.0;
        e = fma(xv[k], k < K ? sw[w][k + 1] : 0.0, e);
      }
      double term = 0.0, r = 0.0;
      if (valid) {
        const int yv = a.y[n];
        stanx::math::sx_logit_row(e, yv ? 1.0 : -1.0, &term, &r);
        if (!isfinite(e)) bad += 1.0;
      }
      s0 += term;
      s1 += r;
#pragma unroll
      for (int k = 0; k < KMAX; ++k) acc[k] = fma(xv[k], r, acc[k]);
    }
    double* out = a.cpart + (ch * a.chains + c) * a.E;
    s1 = wave_sum(s1);
    if (lane == 0) out[0] = s1;
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      if (k < K) {
        const double v = wave_sum(acc[k]);
        if (lane == 0) out[1 + k] = v;
      }
    }
    s0 = wave_sum(s0);
    bad = wave_sum(bad);
    if (lane == 0) {
      out[K + 1] = s0;
      out[K + 2] = bad;
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Bernoulli-logit row term and residual in FP32 for sign s = 2y - 1; s = 0
// (a padding row) gives term = r = 0.
// Hardware transcendentals (v_exp_f32, v_log_f32, v_rcp_f32): absolute
// error of term ~1e-7, a few ulp relative in r.
__device__ __forceinline__ void logit_row_f32s(float e, float s, float* term, float* r) {
  const float se = s * e;
  const float t = __expf(-fabsf(se));
  const float u = 1.0f + t;
  const float inv = __builtin_amdgcn_rcpf(u);
  *term = s == 0.f ? 0.f : -(fmaxf(-se, 0.0f) + __logf(u));
  *r = s * (se >= 0.0f ? t * inv : inv);
}

// FP32, many chains: a wave per chunk, chains on lanes; the chunk's rows
// are staged 64 at a time in the wave's LDS (rows padded to KMAX floats)
// and read as float4 broadcasts.  The next piece's global loads are in
// flight (in registers) while the current piece is computed; rows go four
// at a time, so four independent FMA chains hide the FMA and LDS latency.
// Padding rows (x = 0, sign 0) add nothing.
template <int KMAX>
__global__ __launch_bounds__(256) void hier_pass_f32(hier_args a) {
  static_assert(KMAX % 4 == 0, "rows are read as float4");
  constexpr int PR = 64;  // rows per piece
  __shared__ __attribute__((aligned(16))) float sx[4][PR * KMAX];
  __shared__ float sy[4][PR];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int K = a.K;
  const int c = blockIdx.y * 64 + lane;
  const bool live = c < a.chains;
  const double* th = a.theta + static_cast<int64_t>(live ? c : 0) * a.theta_stride;
  const int64_t ch = static_cast<int64_t>(blockIdx.x) * 4 + w;
  if (ch >= a.n_chunks) return;  // whole wave; no block barrier follows
  const int g = a.chunk_group[ch];
  const int64_t r0 = a.chunk_row0[ch], r1 = a.chunk_row0[ch + 1];
  const int rows = static_cast<int>(r1 - r0);
  const double* z = th + 2 + 2 * K + static_cast<int64_t>(g) * (K + 1);
  float wk[KMAX];
  float b = 0.f;
  {
    const double tau0 = exp(th[1 + K]);
    b = static_cast<float>(th[0] + tau0 * z[0]);
#pragma unroll
    for (int k = 0; k < KMAX; ++k)
      wk[k] = k < K ? static_cast<float>(th[1 + k] + exp(th[2 + K + k]) * z[1 + k]) : 0.f;
  }
  float acc[KMAX];
#pragma unroll
  for (int k = 0; k < KMAX; ++k) acc[k] = 0.f;
  float s1 = 0.f;
  double s0 = 0.0;
  int bad = 0;
  // staging: lane `lane` loads elements lane + 64 j (j < KMAX) of the
  // piece's contiguous rows (coalesced, one base address) and stores them
  // at their padded LDS positions (row * KMAX + column), computed
  // incrementally: adding 64 to an element index adds 64 / K rows and
  // 64 % K columns (with a carry)
  float stg[KMAX];
  float ystg = 0.f;
  int lim = 0;  // elements in the staged piece (cnt * K)
  const int dr = 64 / K, dk = 64 % K;
  const int row0 = lane / K, col0 = lane % K;
#define SX_LOAD_PIECE(P0)                                                        \
  {                                                                              \
    const int cnt_ = rows - (P0) < PR ? rows - (P0) : PR;                        \
    const float* src_ = a.X32 + (r0 + (P0)) * K + lane;                          \
    lim = cnt_ * K;                                                              \
    _Pragma("unroll") for (int j = 0; j < KMAX; ++j)                             \
      stg[j] = lane + 64 * j < lim ? src_[64 * j] : 0.f;                         \
    ystg = lane < cnt_ ? (a.y[r0 + (P0) + lane] ? 1.0f : -1.0f) : 0.f;           \
  }
  SX_LOAD_PIECE(0);
  for (int p0 = 0; p0 < rows; p0 += PR) {
    const int cnt = rows - p0 < PR ? rows - p0 : PR;
    // LDS accesses of one wave complete in order: the previous piece's
    // reads are done before these writes land
    __builtin_amdgcn_wave_barrier();
    // zero the padded piece, then place the staged elements
#pragma unroll
    for (int j = 0; j < KMAX; ++j) sx[w][lane + 64 * j] = 0.f;
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    {
      int row = row0, col = col0;
#pragma unroll
      for (int j = 0; j < KMAX; ++j) {
        if (lane + 64 * j < lim) sx[w][row * KMAX + col] = stg[j];
        row += dr;
        col += dk;
        if (col >= K) {
          col -= K;
          ++row;
        }
      }
    }
    sy[w][lane] = ystg;
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    if (p0 + PR < rows) SX_LOAD_PIECE(p0 + PR);  // in flight during this piece
    float s0p = 0.f;  // the piece's row terms in FP32, added to FP64 per piece
    for (int i = 0; i < cnt; i += 4) {
      float e[4] = {b, b, b, b};
#pragma unroll
      for (int q = 0; q < KMAX / 4; ++q) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float4 xv = *reinterpret_cast<const float4*>(&sx[w][(i + r) * KMAX + 4 * q]);
          e[r] = fmaf(xv.x, wk[4 * q], e[r]);
          e[r] = fmaf(xv.y, wk[4 * q + 1], e[r]);
          e[r] = fmaf(xv.z, wk[4 * q + 2], e[r]);
          e[r] = fmaf(xv.w, wk[4 * q + 3], e[r]);
        }
      }
      float rr[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float sg = sy[w][i + r];
        float term;
        logit_row_f32s(e[r], sg, &term, &rr[r]);
        bad += (sg != 0.f && !isfinite(e[r])) ? 1 : 0;
        s0p += term;
        s1 += rr[r];
      }
      // re-read the rows from LDS for the adjoint rather than keep 4 x KMAX
      // floats live across the link (register pressure sets occupancy)
      asm volatile("" ::: "memory");
#pragma unroll
      for (int q = 0; q < KMAX / 4; ++q) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float4 xv = *reinterpret_cast<const float4*>(&sx[w][(i + r) * KMAX + 4 * q]);
          acc[4 * q] = fmaf(xv.x, rr[r], acc[4 * q]);
          acc[4 * q + 1] = fmaf(xv.y, rr[r], acc[4 * q + 1]);
          acc[4 * q + 2] = fmaf(xv.z, rr[r], acc[4 * q + 2]);
          acc[4 * q + 3] = fmaf(xv.w, rr[r], acc[4 * q + 3]);
        }
      }
    }
    s0 += static_cast<double>(s0p);
  }
#undef SX_LOAD_PIECE
  if (!live) return;
  double* out = a.cpart + (ch * a.chains + c) * a.E;
  out[0] = s1;
#pragma unroll
  for (int k = 0; k < KMAX; ++k)
    if (k < K) out[1 + k] = acc[k];
  out[K + 1] = s0;
  out[K + 2] = static_cast<double>(bad);
}

// Per group and chain: sum the group's chunks in order; z gradients;
// per-group contributions [S_0..S_K, z.S terms (K+1), lp, bad, sum z^2].
// A block takes one group and 8 chains: a chunk's partials of those chains
// are contiguous (chain-major, E entries each), so the block reads them
// coalesced, 32 threads per chain.
__global__ __launch_bounds__(256) void hier_groups(hier_args a) {
  const int g = blockIdx.x;
  const int c = blockIdx.y * 8 + (threadIdx.x >> 5);
  const int e = threadIdx.x & 31;
  if (c >= a.chains) return;
  const int K = a.K;
  const double* th = a.theta + static_cast<int64_t>(c) * a.theta_stride;
  double* out = a.out + static_cast<int64_t>(c) * a.out_stride;
  const int64_t ch0 = a.group_chunk0[g], ch1 = a.group_chunk0[g + 1];
  const double* z = th + 2 + 2 * K + static_cast<int64_t>(g) * (K + 1);
  double* gs = a.gpart + (static_cast<int64_t>(c) * a.J + g) * a.G;
  for (int ee = e; ee < K + 3; ee += 32) {  // K <= 32: at most two rounds
    double v = 0.0;
    for (int64_t ch = ch0; ch < ch1; ++ch) v += a.cpart[(ch * a.chains + c) * a.E + ee];
    if (ee <= K) {
      const double tau = exp(th[1 + K + ee]);
      out[1 + 2 + 2 * K + static_cast<int64_t>(g) * (K + 1) + ee] = tau * v - z[ee];
      gs[ee] = v;
      gs[K + 1 + ee] = tau * z[ee] * v;
    } else {
      gs[K + 1 + ee] = v;  // lp (ee = K+1) at 2K+2, bad (ee = K+2) at 2K+3
    }
  }
  if (e == 0) {
    double zz = 0.0;
    for (int k = 0; k <= K; ++k) zz += z[k] * z[k];
    gs[2 * K + 4] = zz;
  }
}

// Per chain: sums over groups (sixteen fixed interleaved slices of groups,
// each summed with four accumulators, then combined in order), priors,
// Jacobian, output.  Every load of a slice is independent, so the sums are
// bandwidth- rather than latency-bound.
__global__ __launch_bounds__(1024) void hier_final(hier_args a) {
  __shared__ double red[16][64];
  __shared__ double st[128];
  const int c = blockIdx.x;
  const int K = a.K;
  const double* th = a.theta + static_cast<int64_t>(c) * a.theta_stride;
  double* out = a.out + static_cast<int64_t>(c) * a.out_stride;
  const int sl = threadIdx.x >> 6;
  for (int e0 = 0; e0 < a.G; e0 += 64) {
    const int e = e0 + (threadIdx.x & 63);
    if (e < a.G) {
      const double* gp = a.gpart + static_cast<int64_t>(c) * a.J * a.G + e;
      double v[4] = {0.0, 0.0, 0.0, 0.0};
      int g = sl;
      for (; g + 48 < a.J; g += 64) {
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] += gp[static_cast<int64_t>(g + 16 * u) * a.G];
      }
      for (int u = 0; g < a.J; g += 16, ++u) v[u] += gp[static_cast<int64_t>(g) * a.G];
      red[sl][threadIdx.x & 63] = (v[0] + v[1]) + (v[2] + v[3]);
    }
    __syncthreads();
    if (threadIdx.x < 64 && e0 + static_cast<int>(threadIdx.x) < a.G) {
      double t = 0.0;
#pragma unroll
      for (int i = 0; i < 16; ++i) t += red[i][threadIdx.x];
      st[e0 + threadIdx.x] = t;
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double lp = st[2 * K + 2];
  double bad = st[2 * K + 3];
  // alpha
  double ga = st[0];
  if (a.alpha_sd > 0) {
    const double zz = th[0] / a.alpha_sd;
    lp += -0.5 * zz * zz + (a.propto ? 0.0 : a.alpha_const);
    ga += -th[0] / (a.alpha_sd * a.alpha_sd);
  }
  out[1] = ga;
  for (int k = 0; k < K; ++k) {
    double gb = st[1 + k];
    const double bk = th[1 + k];
    if (a.beta_sd > 0) {
      const double zz = bk / a.beta_sd;
      lp += -0.5 * zz * zz + (a.propto ? 0.0 : a.beta_const);
      gb += -bk / (a.beta_sd * a.beta_sd);
    }
    out[2 + k] = gb;
  }
  for (int k = 0; k <= K; ++k) {
    const double v = th[1 + K + k];
    const double tau = exp(v);
    double gv = st[K + 1 + k];
    if (a.tau_sd > 0) {
      const double zz = tau / a.tau_sd;
      lp += -0.5 * zz * zz + (a.propto ? 0.0 : a.tau_const);
      gv += -zz * zz;
    }
    if (a.jacobian) {
      lp += v;
      gv += 1.0;
    }
    out[2 + K + k] = gv;
  }
  lp += -0.5 * st[2 * K + 4];
  if (!a.propto) lp += -static_cast<double>(a.J) * (K + 1) * kHalfLog2PiH;
  out[0] = lp;
  if (!isfinite(lp)) bad += 1.0;
  out[a.D + 1] = bad > 0 ? 1.0 : 0.0;
}

}  // namespace

hipError_t launch_hier(const hier_args& a, bool fp32, hipStream_t s) {
  if (a.K > 32 || 2 * a.K + 5 > 128 || a.chains < 1 || a.J < 1) return hipErrorInvalidValue;
  if (a.n_chunks > 0) {
    if (!fp32) {
      const int64_t want = (a.n_chunks + 3) / 4;
      const dim3 grid(static_cast<unsigned>(want < 2048 ? want : 2048), static_cast<unsigned>(a.chains));
      if (a.K <= 4)
        hipLaunchKernelGGL((hier_pass_f64<4>), grid, dim3(256), 0, s, a);
      else if (a.K <= 8)
        hipLaunchKernelGGL((hier_pass_f64<8>), grid, dim3(256), 0, s, a);
      else if (a.K <= 16)
        hipLaunchKernelGGL((hier_pass_f64<16>), grid, dim3(256), 0, s, a);
      else if (a.K <= 24)
        hipLaunchKernelGGL((hier_pass_f64<24>), grid, dim3(256), 0, s, a);
      else
        hipLaunchKernelGGL((hier_pass_f64<32>), grid, dim3(256), 0, s, a);
    } else {
      const dim3 grid(static_cast<unsigned>((a.n_chunks + 3) / 4), static_cast<unsigned>((a.chains + 63) / 64));
      if (a.K <= 4)
        hipLaunchKernelGGL((hier_pass_f32<4>), grid, dim3(256), 0, s, a);
      else if (a.K <= 8)
        hipLaunchKernelGGL((hier_pass_f32<8>), grid, dim3(256), 0, s, a);
      else if (a.K <= 16)
        hipLaunchKernelGGL((hier_pass_f32<16>), grid, dim3(256), 0, s, a);
      else if (a.K <= 24)
        hipLaunchKernelGGL((hier_pass_f32<24>), grid, dim3(256), 0, s, a);
      else
        hipLaunchKernelGGL((hier_pass_f32<32>), grid, dim3(256), 0, s, a);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(hier_groups, dim3(a.J, (a.chains + 7) / 8), dim3(256), 0, s, a);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(hier_final, dim3(a.chains), dim3(1024), 0, s, a);
  return hipGetLastError();
}

}  // namespace sx
