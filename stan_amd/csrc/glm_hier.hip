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

typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP64: a wave per chunk, rows on lanes; chains on grid.y.  A chunk's rows
// are contiguous in X (row-major), so each block of up to 64 rows is one
// contiguous run of rows * K doubles: the wave loads it with every load
// instruction covering 64 consecutive doubles (512 bytes), the loads of the
// next block in flight while this one is computed, and places it in the
// wave's LDS at row stride K | 1, from where lane r reads row r.  A chunk's
// sums are formed by one wave in a fixed order (lane r takes rows r, r + 64,
// ...; columns in order), so each chain's result depends neither on the
// launch shape nor on the other chains.
template <int KMAX>
__global__ __launch_bounds__(256) void hier_pass_f64(hier_args a) {
  constexpr int BS = 64 * (KMAX | 1);  // LDS doubles of one block
  __shared__ double sx[4][BS + 1];     // + a sink for positions past the block
  __shared__ double sw[4][KMAX + 1];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int K = a.K;
  const int KS = K | 1;
  const int c = blockIdx.y;
  const double* th = a.theta + static_cast<int64_t>(c) * a.theta_stride;
  const int64_t W = static_cast<int64_t>(gridDim.x) * 4;
  // (K = 0, intercepts only: nothing is staged; no division by zero)
  const int Kd = K > 0 ? K : 1;
  const int dr = 64 / Kd, dk = 64 % Kd;
  const int row0 = lane / Kd, col0 = lane % Kd;
  double stg[KMAX];
  int8_t ystg = 0;
  // loads of rows [base, base + 64) of chunk ch into registers; element
  // indices past the block are clamped to its last element (an in-bounds
  // load) and their values never stored
#define SX_LOAD_BLOCK(CH, BASE)                                                  \
  {                                                                              \
    const int64_t e1_ = a.chunk_row0[(CH) + 1];                                  \
    const int cnt_ = static_cast<int>(e1_ - (BASE) < 64 ? e1_ - (BASE) : 64);    \
    const int lim_ = cnt_ * K;                                                   \
    const double* src_ = a.X + (BASE) * K;                                       \
    _Pragma("unroll") for (int j = 0; j < KMAX; ++j) {                           \
      const int e_ = lane + 64 * j;                                              \
      stg[j] = (j < K && lim_ > 0) ? src_[e_ < lim_ ? e_ : lim_ - 1] : 0.0;      \
    }                                                                            \
    ystg = a.y[(BASE) + (lane < cnt_ ? lane : cnt_ - 1)];                        \
  }
  int64_t ch = static_cast<int64_t>(blockIdx.x) * 4 + w;
  if (ch >= a.n_chunks) return;  // whole wave; no block barrier follows
  int64_t base = a.chunk_row0[ch];
  SX_LOAD_BLOCK(ch, base);
  while (ch < a.n_chunks) {
    const int g = a.chunk_group[ch];
    const int64_t r1 = a.chunk_row0[ch + 1];
    const double* z = th + 2 + 2 * K + static_cast<int64_t>(g) * (K + 1);
    if (lane <= K) {
      const double tau = exp(th[1 + K + lane]);
      sw[w][lane] = lane == 0 ? th[0] + tau * z[0] : th[lane] + tau * z[lane];
    }
    double acc[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) acc[k] = 0.0;
    double s0 = 0.0, s1 = 0.0, bad = 0.0;
    for (; base < r1; base += 64) {
      const int cnt = static_cast<int>(r1 - base < 64 ? r1 - base : 64);
      // LDS accesses of one wave complete in order: the previous block's
      // reads are done before these writes land
      __builtin_amdgcn_wave_barrier();
      {
        const int lim = cnt * K;
        int row = row0, col = col0;
#pragma unroll
        for (int j = 0; j < KMAX; ++j) {
          if (j < K) sx[w][lane + 64 * j < lim ? row * KS + col : BS] = stg[j];
          row += dr;
          col += dk;
          if (col >= Kd) {
            col -= Kd;
            ++row;
          }
        }
      }
      const int yv = ystg;
      __builtin_amdgcn_s_waitcnt(0);
      __builtin_amdgcn_wave_barrier();
      // the next block (of this chunk or the wave's next one) in flight
      {
        int64_t nch = ch, nbase = base + 64;
        if (nbase >= r1) {
          nch = ch + W;
          nbase = nch < a.n_chunks ? a.chunk_row0[nch] : 0;
        }
        if (nch < a.n_chunks) SX_LOAD_BLOCK(nch, nbase);
      }
      const bool valid = lane < cnt;
      double e = sw[w][0];
      double xv[KMAX];
#pragma unroll
      for (int k = 0; k < KMAX; ++k) {
        xv[k] = (valid && k < K) ? sx[w][lane * KS + k] : 0.0;
        e = fma(xv[k], k < K ? sw[w][k + 1] : 0.0, e);
      }
      double term = 0.0, r = 0.0;
      if (valid) {
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
    ch += W;
    if (ch < a.n_chunks) base = a.chunk_row0[ch];
  }
#undef SX_LOAD_BLOCK
}

// The row term and residual of the matrix-core pass on the raw hardware
// transcendentals: the exp2 argument is <= 0 and the log2 argument lies in
// [1, 2], where v_exp_f32 and v_log_f32 need none of the range handling of
// the library calls (an exp2 result below 2^-126 may flush to zero; it
// enters u = 1 + t and r = s t / u only at that size).  s = 0 (a padding
// row) gives term = r = 0 through the factor |s|, without a select.
__device__ __forceinline__ void logit_row_lean(float e, float s, float* term, float* r) {
  constexpr float kLog2e = 1.44269504088896341f, kLn2 = 0.693147180559945309f;
  const float se = s * e;
  const float t = __builtin_amdgcn_exp2f(-fabsf(se) * kLog2e);
  const float u = 1.0f + t;
  const float inv = __builtin_amdgcn_rcpf(u);
  *term = -fabsf(s) * fmaf(__builtin_amdgcn_logf(u), kLn2, fmaxf(-se, 0.0f));
  *r = s * (se >= 0.0f ? t * inv : inv);
}

// FP32 on the matrix cores: a wave per chunk and up to 64 chains (NH
// halves of 32).  Within a group every chain's coefficients w (K) and
// intercept b are fixed, so a 32-row tile of the chunk gives two small
// products per half: the forward H (32 rows x 32 chains) = X W and the
// adjoint S^T (32 chains x 32 k) += R^T X, both on
// v_mfma_f32_32x32x2_f32 (FP32 products, FP32 accumulation), with the
// logistic link on the VALU between them.  Operands (lane l, h = l / 32,
// i = l % 32): A[i][h], B[h][i]; result d[j] = D[8 (j/4) + 4 h + j%4][i].
//   forward  step s: A = X[row i][2s + h] (LDS), B = w[2s + h] of chain i
//            (registers, zero past K)
//   adjoint  step s: A = r of chain i at row 8 (s/4) + 4 h + s%4 (the
//            lane's own forward element s), B = X[that row][col i]
// The tile (32 rows x K, contiguous in X) is loaded coalesced into
// registers one tile ahead and placed in the wave's LDS at row stride
// K | 1 (odd: the forward's 32 rows fall in distinct banks).  Rows past
// the chunk have x = 0 and sign 0 and add nothing.
template <int KMAX, int NH>
__global__ __launch_bounds__(256) void hier_mfma_f32(hier_args a) {
  static_assert(KMAX % 4 == 0 && KMAX <= 32, "K <= 32 in buckets of 4");
  static_assert(NH == 1 || NH == 2, "one or two halves of 32 chains");
  constexpr int NJ = KMAX / 2;           // staged floats per lane (32 K / 64), forward steps
  constexpr int TS = 32 * 33 + 1;        // LDS floats of a tile (stride <= 33)
  __shared__ float sx[4][TS + 1];        // + a sink for staging lanes past the tile
  __shared__ float sy[4][32];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int li = lane & 31, h = lane >> 5;
  const int K = a.K;
  const int KS = K | 1;
  const int64_t ch = static_cast<int64_t>(blockIdx.x) * 4 + w;
  if (ch >= a.n_chunks) return;  // whole wave; no block barrier follows
  const int g = a.chunk_group[ch];
  const int64_t r0 = a.chunk_row0[ch], r1 = a.chunk_row0[ch + 1];
  const int rows = static_cast<int>(r1 - r0);
  const int cb = blockIdx.y * (32 * NH);
  // chain coefficients: half hf, chain cb + 32 hf + i, k = 2 s + h
  float wk[NH][NJ], b[NH];
#pragma unroll
  for (int hf = 0; hf < NH; ++hf) {
    const int c = cb + 32 * hf + li;
    const double* th = a.theta + static_cast<int64_t>(c < a.chains ? c : 0) * a.theta_stride;
    const double* z = th + 2 + 2 * K + static_cast<int64_t>(g) * (K + 1);
    b[hf] = static_cast<float>(th[0] + exp(th[1 + K]) * z[0]);
#pragma unroll
    for (int s = 0; s < NJ; ++s) {
      const int k = 2 * s + h;
      wk[hf][s] = k < K ? static_cast<float>(th[1 + k] + exp(th[2 + K + k]) * z[1 + k]) : 0.f;
    }
  }
  f32x16 sacc[NH];
#pragma unroll
  for (int hf = 0; hf < NH; ++hf)
#pragma unroll
    for (int j = 0; j < 16; ++j) sacc[hf][j] = 0.f;
  // nf: sum of 0 * eta, NaN once any eta is infinite or NaN.  A padding
  // row's eta (0 * w + b) is non-finite only if the chain's w or b is, and
  // then so is every real row's
  float s1[NH], s0p[NH], nf[NH];
  double s0[NH];
#pragma unroll
  for (int hf = 0; hf < NH; ++hf) {
    s1[hf] = 0.f;
    nf[hf] = 0.f;
    s0[hf] = 0.0;
  }
  // positions past row 31 or col K - 1 are read (times a zero) but never
  // written: zero them once
  for (int e = lane; e < TS; e += 64) sx[w][e] = 0.f;
  float stg[NJ];
  float ystg = 0.f;
  int lim = 0;
  // (K = 0, intercepts only: nothing is staged; no division by zero)
  const int Kd = K > 0 ? K : 1;
  const int dr = 64 / Kd, dk = 64 % Kd;
  const int row0 = lane / Kd, col0 = lane % Kd;
  // branch-free staging: element indices past the tile's last row are
  // clamped to it (an in-bounds load) and the value replaced by zero
#define SX_LOAD_TILE(T0)                                                         \
  {                                                                              \
    const int cnt_ = rows - (T0) < 32 ? rows - (T0) : 32;                        \
    lim = cnt_ * K;                                                              \
    if (lim > 0) {                                                               \
      const float* src_ = a.X32 + (r0 + (T0)) * K;                               \
      _Pragma("unroll") for (int j = 0; j < NJ; ++j) {                           \
        const int e_ = lane + 64 * j;                                            \
        const float v_ = src_[e_ < lim ? e_ : lim - 1];                          \
        stg[j] = e_ < lim ? v_ : 0.f;                                            \
      }                                                                          \
    }                                                                            \
    const int yi_ = lane < cnt_ ? lane : cnt_ - 1;                               \
    const float yv_ = a.y[r0 + (T0) + yi_] ? 1.0f : -1.0f;                       \
    ystg = lane < cnt_ ? yv_ : 0.f;                                              \
  }
  SX_LOAD_TILE(0);
  for (int t0 = 0; t0 < rows; t0 += 32) {
    // LDS accesses of one wave complete in order: the previous tile's reads
    // are done before these writes land.  All 32 x K positions are written
    // (zeros past the chunk).
    __builtin_amdgcn_wave_barrier();
    {
      int row = row0, col = col0;
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        sx[w][lane + 64 * j < 32 * K ? row * KS + col : TS] = stg[j];
        row += dr;
        col += dk;
        if (col >= K) {
          col -= K;
          ++row;
        }
      }
    }
    if (lane < 32) sy[w][lane] = ystg;
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    if (t0 + 32 < rows) SX_LOAD_TILE(t0 + 32);  // in flight during this tile
    f32x16 hacc[NH];
#pragma unroll
    for (int hf = 0; hf < NH; ++hf)
#pragma unroll
      for (int j = 0; j < 16; ++j) hacc[hf][j] = 0.f;
    // KMAX / 2 steps (columns K .. KMAX - 1 meet w = 0; what they read is
    // the next row's x or the zeroed tail of the buffer, finite either way)
#pragma unroll
    for (int s = 0; s < NJ; ++s) {
      const float xa = sx[w][li * KS + 2 * s + h];
#pragma unroll
      for (int hf = 0; hf < NH; ++hf)
        hacc[hf] = __builtin_amdgcn_mfma_f32_32x32x2f32(xa, wk[hf][s], hacc[hf], 0, 0, 0);
    }
#pragma unroll
    for (int hf = 0; hf < NH; ++hf) {
      s0p[hf] = 0.f;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int row = 8 * (j >> 2) + 4 * h + (j & 3);
        const float sg = sy[w][row];
        const float e = hacc[hf][j] + b[hf];
        float term, rr;
        logit_row_lean(e, sg, &term, &rr);
        nf[hf] = fmaf(e, 0.f, nf[hf]);
        s0p[hf] += term;
        s1[hf] += rr;
        hacc[hf][j] = rr;
      }
      s0[hf] += static_cast<double>(s0p[hf]);
    }
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int row = 8 * (s >> 2) + 4 * h + (s & 3);
      const float xv = sx[w][row * KS + li];
      const float xb = li < K ? xv : 0.f;
#pragma unroll
      for (int hf = 0; hf < NH; ++hf)
        sacc[hf] = __builtin_amdgcn_mfma_f32_32x32x2f32(hacc[hf][s], xb, sacc[hf], 0, 0, 0);
    }
  }
#undef SX_LOAD_TILE
  // sacc[hf][j] = S[chain cb + 32 hf + 8 (j/4) + 4 h + j%4][k = i]; the row
  // terms of chain i are split over the two halves of the wave's rows
#pragma unroll
  for (int hf = 0; hf < NH; ++hf) {
    if (li < K) {
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int c = cb + 32 * hf + 8 * (j >> 2) + 4 * h + (j & 3);
        if (c < a.chains) a.cpart[(ch * a.chains + c) * a.E + 1 + li] = sacc[hf][j];
      }
    }
    const float s1t = s1[hf] + __shfl_xor(s1[hf], 32, 64);
    const double s0t = s0[hf] + __shfl_xor(s0[hf], 32, 64);
    const float nft = nf[hf] + __shfl_xor(nf[hf], 32, 64);
    const int badt = nft == nft ? 0 : 1;
    const int c = cb + 32 * hf + li;
    if (h == 0 && c < a.chains) {
      double* out = a.cpart + (ch * a.chains + c) * a.E;
      out[0] = s1t;
      out[K + 1] = s0t;
      out[K + 2] = static_cast<double>(badt);
    }
  }
}

// Per group and chain: sum the group's chunks in order; z gradients;
// per-group contributions [S_0..S_K, z.S terms (K+1), lp, bad, sum z^2],
// summed over the block's groups in group order: the block's row of
// partials, [(chain x blocks + block) x G].  A block takes kGroupsPerBlock
// consecutive groups and 8 chains (32 threads per chain; tau = exp(v)
// computed once per chain and block): a chunk's partials of those chains
// are contiguous (chain-major, E entries each), so the block reads them
// coalesced.  sum z^2 is a fixed-order shuffle reduction over the chain's
// 32 threads.
constexpr int kGroupsPerBlock = 8;
__global__ __launch_bounds__(256) void hier_groups(hier_args a) {
  const int c = blockIdx.y * 8 + (threadIdx.x >> 5);
  const int e = threadIdx.x & 31;
  if (c >= a.chains) return;  // a whole 32-thread chain group
  const int K = a.K;
  const double* th = a.theta + static_cast<int64_t>(c) * a.theta_stride;
  double* out = a.out + static_cast<int64_t>(c) * a.out_stride;
  // entries e and e + 32 (K <= 32: at most two per thread)
  const double tau0 = e <= K ? exp(th[1 + K + e]) : 0.0;
  const double tau1 = e + 32 <= K ? exp(th[1 + K + e + 32]) : 0.0;
  const int g0 = blockIdx.x * kGroupsPerBlock;
  const int g1 = g0 + kGroupsPerBlock < a.J ? g0 + kGroupsPerBlock : a.J;
  double sv[2] = {0.0, 0.0}, szs[2] = {0.0, 0.0}, szz = 0.0;  // sums over the block's groups
  for (int g = g0; g < g1; ++g) {
    const int64_t ch0 = a.group_chunk0[g], ch1 = a.group_chunk0[g + 1];
    const double* z = th + 2 + 2 * K + static_cast<int64_t>(g) * (K + 1);
    double zz = 0.0;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int ee = e + 32 * r;
      if (ee >= K + 3) continue;
      // chunks in order; eight loads in flight at a time
      double v = 0.0;
      int64_t ch = ch0;
      for (; ch + 8 <= ch1; ch += 8) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = a.cpart[((ch + u) * a.chains + c) * a.E + ee];
#pragma unroll
        for (int u = 0; u < 8; ++u) v += x[u];
      }
      for (; ch < ch1; ++ch) v += a.cpart[(ch * a.chains + c) * a.E + ee];
      sv[r] += v;
      if (ee <= K) {
        const double tau = r == 0 ? tau0 : tau1;
        const double zv = z[ee];
        out[1 + 2 + 2 * K + static_cast<int64_t>(g) * (K + 1) + ee] = tau * v - zv;
        szs[r] += tau * zv * v;
        zz += zv * zv;
      }
    }
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) zz += __shfl_xor(zz, off, 32);
    szz += zz;
  }
  const int nblk = (a.J + kGroupsPerBlock - 1) / kGroupsPerBlock;
  double* gs = a.gpart + (static_cast<int64_t>(c) * nblk + blockIdx.x) * a.G;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int ee = e + 32 * r;
    if (ee >= K + 3) continue;
    if (ee <= K) {
      gs[ee] = sv[r];
      gs[K + 1 + ee] = szs[r];
    } else {
      gs[K + 1 + ee] = sv[r];  // lp (ee = K+1) at 2K+2, bad (ee = K+2) at 2K+3
    }
  }
  if (e == 0) gs[2 * K + 4] = szz;
}

// Per chain: sums over the group blocks' rows (sixteen fixed interleaved
// slices of rows, each summed with four accumulators, then combined in
// order), priors, Jacobian, output.  Every load of a slice is independent
// and sixteen are issued before they are added, so the sums are bandwidth-
// rather than latency-bound.
__global__ __launch_bounds__(1024) void hier_final(hier_args a) {
  __shared__ double red[16][64];
  __shared__ double st[128];
  const int c = blockIdx.x;
  const int K = a.K;
  const double* th = a.theta + static_cast<int64_t>(c) * a.theta_stride;
  double* out = a.out + static_cast<int64_t>(c) * a.out_stride;
  const int sl = threadIdx.x >> 6;
  const int R = (a.J + kGroupsPerBlock - 1) / kGroupsPerBlock;  // rows per chain
  for (int e0 = 0; e0 < a.G; e0 += 64) {
    const int e = e0 + (threadIdx.x & 63);
    if (e < a.G) {
      const double* gp = a.gpart + static_cast<int64_t>(c) * R * a.G + e;
      double v[4] = {0.0, 0.0, 0.0, 0.0};
      int g = sl;
      // sixteen loads in flight, added in the order of the 4-load steps
      for (; g + 240 < R; g += 256) {
        double x[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) x[u] = gp[static_cast<int64_t>(g + 16 * u) * a.G];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u & 3] += x[u];
      }
      for (; g + 48 < R; g += 64) {
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] += gp[static_cast<int64_t>(g + 16 * u) * a.G];
      }
      for (int u = 0; g < R; g += 16, ++u) v[u] += gp[static_cast<int64_t>(g) * a.G];
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
  // prior and Jacobian terms in parallel: wave 0 lane 0 alpha, lanes
  // 1..K beta; wave 1 lanes 0..K tau.  Thread 0 then adds the log-density
  // terms in a fixed order (alpha, beta_k, then per tau_k the prior and the
  // Jacobian term).
  __shared__ double lpa[33], lpt[2][33];
  const int ln = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (wv == 0 && ln <= K) {
    if (ln == 0) {
      double ga = st[0], t = 0.0;
      if (a.alpha_sd > 0) {
        const double zz = th[0] / a.alpha_sd;
        t = -0.5 * zz * zz + (a.propto ? 0.0 : a.alpha_const);
        ga += -th[0] / (a.alpha_sd * a.alpha_sd);
      }
      out[1] = ga;
      lpa[0] = t;
    } else {
      const int k = ln - 1;
      double gb = st[1 + k], t = 0.0;
      const double bk = th[1 + k];
      if (a.beta_sd > 0) {
        const double zz = bk / a.beta_sd;
        t = -0.5 * zz * zz + (a.propto ? 0.0 : a.beta_const);
        gb += -bk / (a.beta_sd * a.beta_sd);
      }
      out[2 + k] = gb;
      lpa[ln] = t;
    }
  } else if (wv == 1 && ln <= K) {
    const int k = ln;
    const double v = th[1 + K + k];
    const double tau = exp(v);
    double gv = st[K + 1 + k], t = 0.0, jt = 0.0;
    if (a.tau_sd > 0) {
      const double zz = tau / a.tau_sd;
      t = -0.5 * zz * zz + (a.propto ? 0.0 : a.tau_const);
      gv += -zz * zz;
    }
    if (a.jacobian) {
      jt = v;
      gv += 1.0;
    }
    out[2 + K + k] = gv;
    lpt[0][k] = t;
    lpt[1][k] = jt;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double lp = st[2 * K + 2];
  double bad = st[2 * K + 3];
  if (a.alpha_sd > 0) lp += lpa[0];
  if (a.beta_sd > 0)
    for (int k = 0; k < K; ++k) lp += lpa[1 + k];
  for (int k = 0; k <= K; ++k) {
    if (a.tau_sd > 0) lp += lpt[0][k];
    if (a.jacobian) lp += lpt[1][k];
  }
  lp += -0.5 * st[2 * K + 4];
  if (!a.propto) lp += -static_cast<double>(a.J) * (K + 1) * kHalfLog2PiH;
  out[0] = lp;
  if (!isfinite(lp)) bad += 1.0;
  out[a.D + 1] = bad > 0 ? 1.0 : 0.0;
}

template <int NH>
hipError_t launch_hier_mfma(const hier_args& a, dim3 grid, hipStream_t s) {
  switch ((a.K + 3) / 4) {
    case 0:
    case 1:
    case 2: hipLaunchKernelGGL((hier_mfma_f32<8, NH>), grid, dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL((hier_mfma_f32<12, NH>), grid, dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL((hier_mfma_f32<16, NH>), grid, dim3(256), 0, s, a); break;
    case 5: hipLaunchKernelGGL((hier_mfma_f32<20, NH>), grid, dim3(256), 0, s, a); break;
    case 6: hipLaunchKernelGGL((hier_mfma_f32<24, NH>), grid, dim3(256), 0, s, a); break;
    case 7: hipLaunchKernelGGL((hier_mfma_f32<28, NH>), grid, dim3(256), 0, s, a); break;
    case 8: hipLaunchKernelGGL((hier_mfma_f32<32, NH>), grid, dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
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
      const dim3 grid(static_cast<unsigned>((a.n_chunks + 3) / 4), static_cast<unsigned>((a.chains + 31) / 32));
      const hipError_t e = launch_hier_mfma<1>(a, grid, s);
      if (e != hipSuccess) return e;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(hier_groups, dim3((a.J + kGroupsPerBlock - 1) / kGroupsPerBlock, (a.chains + 7) / 8),
                     dim3(256), 0, s, a);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(hier_final, dim3(a.chains), dim3(1024), 0, s, a);
  return hipGetLastError();
}

}  // namespace sx
