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
#include "mfma_tile.hpp"
#include "sx_internal.hpp"

namespace sx {

namespace {

constexpr double kHalfLog2PiH = 0.91893853320467274178032973640562;

// FP64: a wave per chunk, rows on lanes; chains on grid.y.  A chunk's rows
// are contiguous in X (row-major), so each block of up to 64 rows is one
// contiguous run of rows * K doubles: the wave loads it with every load
// instruction covering 64 consecutive doubles (512 bytes), the loads of the
// next block in flight while this one is computed, and places it in the
// wave's LDS at row stride K | 1, from where lane r reads row r.  A chunk's
// sums are formed by one wave in a fixed order (lane r takes rows r, r + 64,
// ...; columns in order), so each chain's result depends neither on the
// launch shape nor on the other chains.
template <int K>
__global__ __launch_bounds__(256) void hier_pass_f64(hier_args a) {
  static_assert(K >= 0 && K <= 32, "K <= 32");
  constexpr int KS = K | 1;
  constexpr int BS = 64 * KS;           // LDS doubles of one block
  constexpr int Kd = K > 0 ? K : 1;
  constexpr int DR = 64 / Kd, DK = 64 % Kd;
  __shared__ double sx[4][BS + 1];      // + a sink for positions past the block
  __shared__ double sw[4][K + 1];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = blockIdx.y;
  const double* th = a.theta + static_cast<int64_t>(c) * a.theta_stride;
  const int64_t W = static_cast<int64_t>(gridDim.x) * 4;
  const int row0 = lane / Kd, col0 = lane % Kd;
  double stg[K > 0 ? K : 1];
  int8_t ystg = 0;
  // loads of rows [base, base + 64) of chunk ch into registers: always 64 K
  // consecutive doubles (X carries 64 K zero doubles past its last row, so
  // a block past a chunk's end stays in bounds); the values past the
  // chunk's rows are never stored
#define SX_LOAD_BLOCK(CH, BASE)                                                  \
  {                                                                              \
    const int64_t e1_ = a.chunk_row0[(CH) + 1];                                  \
    const int cnt_ = static_cast<int>(e1_ - (BASE) < 64 ? e1_ - (BASE) : 64);    \
    const double* src_ = a.X + (BASE) * K + lane;                                \
    _Pragma("unroll") for (int j = 0; j < K; ++j) stg[j] = src_[64 * j];         \
    ystg = a.y[(BASE) + (lane < cnt_ ? lane : cnt_ - 1)];                        \
  }
  int64_t ch = static_cast<int64_t>(blockIdx.x) * 4 + w;
  if (ch >= a.n_chunks) return;  // whole wave; no block barrier follows
  // positions a block leaves unwritten are read by the rows past a chunk's
  // end (their residual is 0): they must hold finite values from the start
  for (int e = lane; e <= BS; e += 64) sx[w][e] = 0.0;
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
    double acc[K > 0 ? K : 1];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
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
        for (int j = 0; j < K; ++j) {
          sx[w][lane + 64 * j < lim ? row * KS + col : BS] = stg[j];
          row += DR;
          col += DK;
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
      // lane r: row r of the block (columns in order); rows past the chunk
      // read earlier blocks' rows or zeros (finite) and get r = 0
      double e = sw[w][0];
#pragma unroll
      for (int k = 0; k < K; ++k) e = fma(sx[w][lane * KS + k], sw[w][k + 1], e);
      double term = 0.0, r = 0.0;
      if (valid) {
        stanx::math::sx_logit_row(e, yv ? 1.0 : -1.0, &term, &r);
        if (!isfinite(e)) bad += 1.0;
      }
      s0 += term;
      s1 += r;
      // read the row again rather than hold it through the link
      asm volatile("" ::: "memory");
#pragma unroll
      for (int k = 0; k < K; ++k) acc[k] = fma(sx[w][lane * KS + k], r, acc[k]);
    }
    double* out = a.cpart + (ch * a.chains + c) * a.E;
    s1 = wave_sum(s1);
    if (lane == 0) out[0] = s1;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const double v = wave_sum(acc[k]);
      if (lane == 0) out[1 + k] = v;
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

// FP32 on the matrix cores: a wave per chunk and up to 64 chains (NH
// halves of 32).  X is stored sign-folded, x' = s x with s = 2y - 1, and
// each staged row carries s itself as column K, so one product per tile
// and half gives the base-2 margin of every (row, chain):
//   q = log2(e) s eta = sum_k x'_k w'_k + s b',  w' = log2(e) w, b' = log2(e) b
// (w, b: the chain's coefficients and intercept in the chunk's group,
// scaled by log2(e) in FP64 and rounded once).  With L = log2(1 + 2^-|q|)
// the row's log density is -ln2 (L + max(-q, 0)) and its scaled residual
// r' = sigma(-s eta) = 2^-(L + max(q, 0)), so the link costs three
// hardware transcendentals (v_exp_f32, v_log_f32, v_exp_f32) and a few
// adds.  The adjoint product sum_rows r' [x' | s] then yields the
// gradient sums S_k = sum r x_k (columns k < K, since r = s r' and s^2 =
// 1) and S_0 = sum r (column K; for K = 32, past the product's 32
// columns, a VALU sum of r' s).  Both products are
// v_mfma_f32_32x32x2_f32 (FP32 products and accumulation).  Operands
// (lane l, h = l / 32, i = l % 32): A[i][h], B[h][i]; result d[j] =
// D[8 (j/4) + 4 h + j%4][i].
//   forward  step s: A = [x'|s][row i][2s + h] (LDS), B = w'[2s + h] of
//            chain i (registers; b' at column K, zero past it)
//   adjoint  step s: A = r' of chain i at row 8 (s/4) + 4 h + s%4 (the
//            lane's own forward element s), B = [x'|s][that row][col i]
//            (columns past K read whatever follows in LDS, always finite,
//            and are never stored)
// The tile (32 rows x K, contiguous in X) is loaded coalesced into
// registers one tile ahead and placed in the wave's LDS at row stride
// (K + 1) | 1 (odd: the forward's 32 rows fall in distinct banks).  Rows
// past the chunk have x' = 0 and s = 0, so q = 0 exactly: their r' adds
// nothing to the adjoint, and their log-density value L(0), computed once
// by the same instruction, is subtracted per padding row.
template <int K, int NH>
__global__ __launch_bounds__(256) void hier_mfma_f32(hier_args a) {
  static_assert(K >= 0 && K <= 32, "K <= 32");
  static_assert(NH == 1 || NH == 2, "one or two halves of 32 chains");
  constexpr int KS = (K + 1) | 1;          // LDS row stride
  constexpr int NJ = (K + 2) / 2;          // forward steps: columns 0 .. 2 NJ - 1 >= K
  constexpr int NS = (32 * K + 63) / 64;   // staged floats per lane
  constexpr int TS = 31 * KS + 34;         // covers every forward and adjoint read
  constexpr float kLog2e = 1.44269504088896341f;
  __shared__ float sx[4][TS + 1];          // + a sink for staging lanes past the tile
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int li = lane & 31, h = lane >> 5;
  const int64_t ch = static_cast<int64_t>(blockIdx.x) * 4 + w;
  if (ch >= a.n_chunks) return;  // whole wave; no block barrier follows
  const int g = a.chunk_group[ch];
  const int64_t r0 = a.chunk_row0[ch], r1 = a.chunk_row0[ch + 1];
  const int rows = static_cast<int>(r1 - r0);
  const int cb = blockIdx.y * (32 * NH);
  // chain coefficients: half hf, chain cb + 32 hf + i, column 2 s + h
  float wk[NH][NJ];
#pragma unroll
  for (int hf = 0; hf < NH; ++hf) {
    const int c = cb + 32 * hf + li;
    const double* th = a.theta + static_cast<int64_t>(c < a.chains ? c : 0) * a.theta_stride;
    const double* z = th + 2 + 2 * K + static_cast<int64_t>(g) * (K + 1);
#pragma unroll
    for (int s = 0; s < NJ; ++s) {
      const int k = 2 * s + h;
      double v = 0.0;
      if (k < K)
        v = th[1 + k] + exp(th[2 + K + k]) * z[1 + k];
      else if (k == K)
        v = th[0] + exp(th[1 + K]) * z[0];
      wk[hf][s] = static_cast<float>(v * 1.4426950408889634074);
    }
  }
  f32x16 sacc[NH];
#pragma unroll
  for (int hf = 0; hf < NH; ++hf)
#pragma unroll
    for (int j = 0; j < 16; ++j) sacc[hf][j] = 0.f;
  // s0: the lane's rows' log-density sums in base 2 (L + max(-q, 0)).
  // They also carry the non-finite flag: |q| - q is NaN or +inf exactly
  // when q is, so a non-finite q makes the sum non-finite.
  double s0[NH];
  // K = 32: the sign column lies past the adjoint's 32 output columns;
  // sum r' s (= sum r) on the VALU instead, per lane over its rows
  float s1v[NH];
#pragma unroll
  for (int hf = 0; hf < NH; ++hf) {
    s0[hf] = 0.0;
    s1v[hf] = 0.f;
  }
  const float lpad = __builtin_amdgcn_logf(1.0f + __builtin_amdgcn_exp2f(-0.0f));
  // positions past row 31 or col K are read (times a zero) but never
  // written: zero them once
  for (int e = lane; e <= TS; e += 64) sx[w][e] = 0.f;
  // LDS position of each staged float (row e / K, column e % K of the
  // tile); the sink past the tile for e >= 32 K
  constexpr int Kd = K > 0 ? K : 1;
  int dst[NS > 0 ? NS : 1];
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const int e = lane + 64 * j;
    dst[j] = e < 32 * K ? (e / Kd) * KS + e % Kd : TS;
  }
  // staging registers: the next tile's floats, its row count and the raw
  // y byte of the lane's row.  Loads stay in flight through a whole tile;
  // nothing reads them before the next tile's LDS writes.  A tile always
  // loads 64 NS consecutive floats: X carries 32 K + 64 zero floats past
  // its last row, so the reads past a chunk's end stay in bounds, and the
  // values past the tile's rows are replaced by zero when written.
  float stg[NS > 0 ? NS : 1];
  int ybyte = 0, scnt = 0;
#define SX_LOAD_TILE(T0)                                                          \
  {                                                                               \
    scnt = rows - (T0) < 32 ? rows - (T0) : 32;                                   \
    const float* src_ = a.X32 + (r0 + (T0)) * K + lane;                           \
    _Pragma("unroll") for (int j = 0; j < NS; ++j) stg[j] = src_[64 * j];         \
    ybyte = a.y[r0 + (T0) + (lane < scnt ? lane : scnt - 1)];                     \
  }
  SX_LOAD_TILE(0);
  for (int t0 = 0; t0 < rows; t0 += 32) {
    const int cnt = scnt;
    // LDS accesses of one wave complete in order: the previous tile's reads
    // are done before these writes land.  All 32 x (K + 1) positions are
    // written (zeros past the chunk).
    __builtin_amdgcn_wave_barrier();
    if (cnt == 32) {
#pragma unroll
      for (int j = 0; j < NS; ++j) sx[w][dst[j]] = stg[j];
    } else {
#pragma unroll
      for (int j = 0; j < NS; ++j) sx[w][dst[j]] = lane + 64 * j < cnt * K ? stg[j] : 0.f;
    }
    if (lane < 32) sx[w][lane * KS + K] = lane < cnt ? (ybyte ? 1.0f : -1.0f) : 0.f;
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    if (t0 + 32 < rows) SX_LOAD_TILE(t0 + 32);  // in flight during this tile
    float xa[NJ];
#pragma unroll
    for (int s = 0; s < NJ; ++s) xa[s] = sx[w][li * KS + 2 * s + h];
    f32x16 hacc[NH];
#pragma unroll
    for (int hf = 0; hf < NH; ++hf)
#pragma unroll
      for (int j = 0; j < 16; ++j) hacc[hf][j] = 0.f;
#pragma unroll
    for (int s = 0; s < NJ; ++s)
#pragma unroll
      for (int hf = 0; hf < NH; ++hf)
        hacc[hf] = __builtin_amdgcn_mfma_f32_32x32x2f32(xa[s], wk[hf][s], hacc[hf], 0, 0, 0);
    float xb[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) xb[s] = sx[w][(8 * (s >> 2) + 4 * h + (s & 3)) * KS + li];
    // padding rows among the lane's 16 (rows 8 b + 4 h + 0..3)
    int npad = 0;
    if (cnt < 32) {
#pragma unroll
      for (int b4 = 0; b4 < 4; ++b4) {
        const int p = 8 * b4 + 4 * h + 4 - cnt;
        npad += p < 0 ? 0 : (p > 4 ? 4 : p);
      }
    }
    float sg[16];
    if constexpr (K == 32) {
#pragma unroll
      for (int j = 0; j < 16; ++j) sg[j] = sx[w][(8 * (j >> 2) + 4 * h + (j & 3)) * KS + K];
    }
    // per (row, chain): L = log2(1 + 2^-|q|); log density (base 2, negated)
    // L + max(-q, 0) = L + (|q| - q) / 2; r' = 2^-(L + (q + |q|) / 2).
    // Both halves are exact.
#pragma unroll
    for (int hf = 0; hf < NH; ++hf) {
      float al = 0.f, am = 0.f;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const float q = hacc[hf][j];
        const float L = __builtin_amdgcn_logf(1.0f + __builtin_amdgcn_exp2f(-fabsf(q)));
        al += L;
        am += fabsf(q) - q;
        hacc[hf][j] = __builtin_amdgcn_exp2f(-fmaf(0.5f, q + fabsf(q), L));
        if constexpr (K == 32) s1v[hf] = fmaf(hacc[hf][j], sg[j], s1v[hf]);
      }
      s0[hf] += static_cast<double>(fmaf(0.5f, am, al)) - static_cast<double>(npad) * lpad;
    }
#pragma unroll
    for (int s = 0; s < 16; ++s)
#pragma unroll
      for (int hf = 0; hf < NH; ++hf)
        sacc[hf] = __builtin_amdgcn_mfma_f32_32x32x2f32(hacc[hf][s], xb[s], sacc[hf], 0, 0, 0);
  }
#undef SX_LOAD_TILE
  // sacc[hf][j] = column i of chain cb + 32 hf + 8 (j/4) + 4 h + j%4:
  // sum r x_i for i < K (slot 1 + i), sum r for i = K (slot 0).  The
  // log-density sums of chain i are split over the two halves of the
  // wave's rows.
#pragma unroll
  for (int hf = 0; hf < NH; ++hf) {
    if (li <= K) {
      const int slot = li < K ? 1 + li : 0;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int c = cb + 32 * hf + 8 * (j >> 2) + 4 * h + (j & 3);
        if (c < a.chains) a.cpart[(ch * a.chains + c) * a.E + slot] = sacc[hf][j];
      }
    }
    const double s0t = s0[hf] + __shfl_xor(s0[hf], 32, 64);
    const int badt = isfinite(s0t) ? 0 : 1;
    const float s1t = s1v[hf] + __shfl_xor(s1v[hf], 32, 64);
    const int c = cb + 32 * hf + li;
    if (h == 0 && c < a.chains) {
      double* out = a.cpart + (ch * a.chains + c) * a.E;
      if constexpr (K == 32) out[0] = s1t;
      out[K + 1] = -0.69314718055994530942 * s0t;
      out[K + 2] = static_cast<double>(badt);
    }
  }
}

// x (FP32) = h + m + l exactly in three BF16 terms, returned as FP32 bit
// patterns whose low halves are zero: h keeps the leading 8 significant
// bits (truncated), m the next 8 of the exact residual x - h, l the exact
// rest (at most 8 bits).  |m| < 2^-7 |h|, |l| < 2^-7 |m|.
__device__ __forceinline__ void split3(float x, uint32_t& h, uint32_t& m, uint32_t& l) {
  h = __float_as_uint(x) & 0xffff0000u;
  const float r1 = x - __uint_as_float(h);
  m = __float_as_uint(r1) & 0xffff0000u;
  l = __float_as_uint(r1 - __uint_as_float(m));
}

// two split terms (high halves) packed as BF16 elements 2i, 2i + 1
__device__ __forceinline__ uint32_t pack_hi(uint32_t lo_elem, uint32_t hi_elem) {
  return (hi_elem & 0xffff0000u) | (lo_elem >> 16);
}

// The x6 pass's chain coefficients, once per launch: for group g and chain
// c, column k of w' = log2(e) (beta_k + tau_k z_{g,k+1}) (k < K) and
// log2(e) (alpha + tau_0 z_{g,0}) (k = K), tau = exp(v), split into three
// BF16 terms (split3 of the FP32 value), stored at
// wsplit[((g C + c) 3 + term) 32 + k]; columns above K stay zero (cleared
// at allocation).  One thread per (column k <= K, chain, 8 groups): tau is
// computed once per thread.
constexpr int kCoefGroups = 8;
__global__ __launch_bounds__(256) void hier_coef_x6(hier_args a) {
  const int64_t id = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int k = static_cast<int>(id & 31);
  const int64_t cg = id >> 5;  // group block * chains + c
  const int K = a.K;
  const int c = static_cast<int>(cg % a.chains);
  const int g0 = static_cast<int>(cg / a.chains) * kCoefGroups;
  if (k > K || g0 >= a.J) return;
  const double* th = a.theta + static_cast<int64_t>(c) * a.theta_stride;
  const double base = k < K ? th[1 + k] : th[0];
  const double tau = k < K ? exp(th[2 + K + k]) : exp(th[1 + K]);
  const double* z = th + 2 + 2 * K + static_cast<int64_t>(g0) * (K + 1) + (k < K ? 1 + k : 0);
  double zv[kCoefGroups];  // every load in flight before the first use
#pragma unroll
  for (int u = 0; u < kCoefGroups; ++u) zv[u] = g0 + u < a.J ? z[u * (K + 1)] : 0.0;
#pragma unroll
  for (int u = 0; u < kCoefGroups; ++u) {
    if (g0 + u >= a.J) break;
    const double v = base + tau * zv[u];
    uint32_t hh, mm, ll;
    split3(static_cast<float>(v * 1.4426950408889634074), hh, mm, ll);
    uint16_t* w = a.wsplit + (static_cast<int64_t>(g0 + u) * a.chains + c) * 96 + k;
    w[0] = static_cast<uint16_t>(hh >> 16);
    w[32] = static_cast<uint16_t>(mm >> 16);
    w[64] = static_cast<uint16_t>(ll >> 16);
  }
}

// The same pass with FP32 products formed on the BF16 matrix cores
// (K <= 31).  Every FP32 operand is split exactly into three BF16 terms,
// x' = xh + xm + xl (X, split once at set-up), w' = wh + wm + wl (the
// chain coefficients), r' = rh + rm + rl (the residuals), and each product
// keeps the six cross terms down to 2^-14 of the leading one:
//   forward  Q += Xh Wh + Xh Wm + Xm Wh + Xh Wl + Xl Wh + Xm Wm
//   adjoint  S += Rh Xh + Rh Xm + Rm Xh + Rh Xl + Rl Xh + Rm Xm
// (the three dropped terms are below 2^-20 of the product), with FP32
// accumulation on v_mfma_f32_32x32x16_bf16.  The BF16
// matrix cores issue beside the VALU link of other waves, which the FP32
// 32x32x2 instruction does not.  Tile layout: 32 rows x 32 columns (x'
// in columns < K, the sign s in column K, zeros past it) per term, in the
// wave's LDS with the 16-byte-chunk XOR permutation of the batched GLM
// pass; the forward operand (8 consecutive columns of a row) is one
// 16-byte read per term, the adjoint's (8 rows of the lane's column,
// matching the rows of the lane's own residuals) two transposing reads.
// The terms of X (sign column included) come from the split copy XS, NCH
// 16-byte chunks per row and term (columns 0 .. 8 NCH - 1), so a tile is
// 32 x 3 NCH contiguous chunks that the wave copies to their permuted LDS
// positions (16-byte loads and stores, no VALU split); rows past the
// chunk's end are stored as zeros.  Link, padding rows and outputs as in
// hier_mfma_f32.
template <int K>
__global__ __launch_bounds__(256) void hier_mfma_x6(hier_args a) {
  static_assert(K >= 0 && K <= 31, "K + 1 <= 32 columns");
  constexpr int NCH = hier_x6_chunks(K);   // 16-byte chunks per row and term
  constexpr int RC = 3 * NCH;              // chunks per row
  constexpr int TC = 32 * RC;              // chunks per tile
  constexpr int NL = (TC + 63) / 64;       // chunk copies per lane and tile
  constexpr int KST = (K + 16) / 16;       // forward k-steps of 16 (columns 0 .. K)
  constexpr int ROWB = 64;                 // bytes of a tile row (32 BF16)
  constexpr int TB = 32 * ROWB;            // bytes of one term's tile
  constexpr int WB = 3 * TB;               // bytes per wave
  __shared__ __attribute__((aligned(16))) char lds[4][WB];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int li = lane & 31, h = lane >> 5;
  // block -> (group of 4 chunks, block of 32 chains): the chain blocks of a
  // chunk group are 8 block ids apart, i.e. on one XCD and dispatched
  // together, so all but the first read the chunk's tiles from that XCD's
  // L2 (launch_hier_x6: a 1-D grid of chunk groups rounded up to 8 x chain
  // blocks)
  const int ncb = (a.chains + 31) / 32;
  const int bq = static_cast<int>(blockIdx.x) >> 3;
  const int cbi = bq % ncb;
  const int64_t ch = (static_cast<int64_t>(bq / ncb) * 8 + (blockIdx.x & 7)) * 4 + w;
  if (ch >= a.n_chunks) return;  // whole wave; no block barrier follows
  const int g = a.chunk_group[ch];
  const int64_t r0 = a.chunk_row0[ch], r1 = a.chunk_row0[ch + 1];
  const int rows = static_cast<int>(r1 - r0);
  const int cb = cbi * 32;
  char* const base = lds[w];
  // chain coefficients (w' = log2(e) w, b' = log2(e) b at column K), split
  // by hier_coef_x6: the B operand of k-step s is columns 16 s + 8 h .. + 7
  // of chain cb + i, one 16-byte load per term
  bf16x8 bwh[KST], bwm[KST], bwl[KST];
  {
    const int c = cb + li < a.chains ? cb + li : 0;
    const uint4* wq =
        reinterpret_cast<const uint4*>(a.wsplit) + (static_cast<int64_t>(g) * a.chains + c) * 12;
#pragma unroll
    for (int s = 0; s < KST; ++s) {
      bwh[s] = __builtin_bit_cast(bf16x8, wq[2 * s + h]);
      bwm[s] = __builtin_bit_cast(bf16x8, wq[4 + 2 * s + h]);
      bwl[s] = __builtin_bit_cast(bf16x8, wq[8 + 2 * s + h]);
    }
  }
  f32x16 sacc;
#pragma unroll
  for (int j = 0; j < 16; ++j) sacc[j] = 0.f;
  double s0 = 0.0;
  const float lpad = __builtin_amdgcn_logf(1.0f + __builtin_amdgcn_exp2f(-0.0f));
  // all three terms zero: columns past K and rows past the chunk stay zero
  for (int e = lane; e < WB / 16; e += 64)
    *reinterpret_cast<uint4*>(base + 16 * e) = make_uint4(0u, 0u, 0u, 0u);
  // LDS byte offset and tile row of the chunks this lane copies: chunk
  // e = lane + 64 j of a tile is chunk e % NCH of term (e / NCH) % 3 of row
  // e / RC; the same for every tile
  int dst[NL], drow[NL];
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    const int e = lane + 64 * j < TC ? lane + 64 * j : 0;  // past TC: not stored
    dst[j] = split_chunk_offset<ROWB, 4, NCH>(e, TB);
    drow[j] = lane + 64 * j < TC ? split_chunk_row<NCH>(e) : 32;
  }
  const uint4* const xs = reinterpret_cast<const uint4*>(a.XS);
  uint4 stg[NL];
  int scnt = 0;
#define SX_LOAD_TILE(T0)                                                          \
  {                                                                               \
    scnt = rows - (T0) < 32 ? rows - (T0) : 32;                                   \
    const uint4* src_ = xs + (r0 + (T0)) * RC + lane;                             \
    _Pragma("unroll") for (int j = 0; j < NL; ++j)                                \
      if (64 * (j + 1) <= TC || lane + 64 * j < TC) stg[j] = src_[64 * j];        \
  }
  SX_LOAD_TILE(0);
  int ln = lane;
  asm volatile("" : "+v"(ln));
  const tile_addr<ROWB, 4> ad(ln);
  for (int t0 = 0; t0 < rows; t0 += 32) {
    const int cnt = scnt;
    // LDS accesses of one wave complete in order: the previous tile's reads
    // are done before these writes land
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int j = 0; j < NL; ++j) {
      if (64 * (j + 1) <= TC || lane + 64 * j < TC)
        *reinterpret_cast<uint4*>(base + dst[j]) =
            drow[j] < cnt ? stg[j] : make_uint4(0u, 0u, 0u, 0u);
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    if (t0 + 32 < rows) SX_LOAD_TILE(t0 + 32);  // in flight during this tile
    f32x16 hacc;
#pragma unroll
    for (int j = 0; j < 16; ++j) hacc[j] = 0.f;
#pragma unroll
    for (int s = 0; s < KST; ++s) {
      const bf16x8 xh = *reinterpret_cast<const bf16x8*>(base + ad.fwd(s));
      const bf16x8 xm = *reinterpret_cast<const bf16x8*>(base + TB + ad.fwd(s));
      const bf16x8 xl = *reinterpret_cast<const bf16x8*>(base + 2 * TB + ad.fwd(s));
      hacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, bwh[s], hacc, 0, 0, 0);
      hacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, bwm[s], hacc, 0, 0, 0);
      hacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xm, bwh[s], hacc, 0, 0, 0);
      hacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, bwl[s], hacc, 0, 0, 0);
      hacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xl, bwh[s], hacc, 0, 0, 0);
      hacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xm, bwm[s], hacc, 0, 0, 0);
    }
    int npad = 0;
    if (cnt < 32) {
#pragma unroll
      for (int b4 = 0; b4 < 4; ++b4) {
        const int p = 8 * b4 + 4 * h + 4 - cnt;
        npad += p < 0 ? 0 : (p > 4 ? 4 : p);
      }
    }
    float al = 0.f, am = 0.f;
#pragma unroll
    for (int st = 0; st < 2; ++st) {
      uint32_t ph[8], pm[8], pl[8];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const float q = hacc[8 * st + jj];
        const float L = __builtin_amdgcn_logf(1.0f + __builtin_amdgcn_exp2f(-fabsf(q)));
        al += L;
        am += fabsf(q) - q;
        split3(__builtin_amdgcn_exp2f(-fmaf(0.5f, q + fabsf(q), L)), ph[jj], pm[jj], pl[jj]);
      }
      uint32_t uh[4], um[4], ul[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        uh[i] = pack_hi(ph[2 * i], ph[2 * i + 1]);
        um[i] = pack_hi(pm[2 * i], pm[2 * i + 1]);
        ul[i] = pack_hi(pl[2 * i], pl[2 * i + 1]);
      }
      const bf16x8 rh = __builtin_bit_cast(bf16x8, uh);
      const bf16x8 rm = __builtin_bit_cast(bf16x8, um);
      const bf16x8 rl = __builtin_bit_cast(bf16x8, ul);
      const bf16x8 th = tr_operand(base, ad.tr(st, 0, 0), ad.tr(st, 1, 0));
      const bf16x8 tm = tr_operand(base + TB, ad.tr(st, 0, 0), ad.tr(st, 1, 0));
      const bf16x8 tl = tr_operand(base + 2 * TB, ad.tr(st, 0, 0), ad.tr(st, 1, 0));
      sacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh, th, sacc, 0, 0, 0);
      sacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh, tm, sacc, 0, 0, 0);
      sacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rm, th, sacc, 0, 0, 0);
      sacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh, tl, sacc, 0, 0, 0);
      sacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rl, th, sacc, 0, 0, 0);
      sacc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rm, tm, sacc, 0, 0, 0);
    }
    s0 += static_cast<double>(fmaf(0.5f, am, al)) - static_cast<double>(npad) * lpad;
  }
#undef SX_LOAD_TILE
  // sacc[j] = column i of chain cb + 8 (j/4) + 4 h + j%4: sum r x_i for
  // i < K (slot 1 + i), sum r for i = K (slot 0)
  if (li <= K) {
    const int slot = li < K ? 1 + li : 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int c = cb + 8 * (j >> 2) + 4 * h + (j & 3);
      if (c < a.chains) a.cpart[(ch * a.chains + c) * a.E + slot] = sacc[j];
    }
  }
  const double s0t = s0 + __shfl_xor(s0, 32, 64);
  const int c = cb + li;
  if (h == 0 && c < a.chains) {
    double* out = a.cpart + (ch * a.chains + c) * a.E;
    out[K + 1] = -0.69314718055994530942 * s0t;
    out[K + 2] = isfinite(s0t) ? 0.0 : 1.0;
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
template <int NR>  // entries per thread: 1 for K + 3 <= 32, else 2
__global__ __launch_bounds__(256) void hier_groups(hier_args a) {
  const int c = blockIdx.y * 8 + (threadIdx.x >> 5);
  const int e = threadIdx.x & 31;
  if (c >= a.chains) return;  // a whole 32-thread chain group
  const int K = a.K;
  const double* th = a.theta + static_cast<int64_t>(c) * a.theta_stride;
  double* out = a.out + static_cast<int64_t>(c) * a.out_stride;
  // entries e and e + 32 (K <= 32: at most two per thread)
  const double tau0 = e <= K ? exp(th[1 + K + e]) : 0.0;
  const double tau1 = NR > 1 && e + 32 <= K ? exp(th[1 + K + e + 32]) : 0.0;
  const int g0 = blockIdx.x * kGroupsPerBlock;
  const int g1 = g0 + kGroupsPerBlock < a.J ? g0 + kGroupsPerBlock : a.J;
  double sv[NR], szs[NR], szz = 0.0;  // sums over the block's groups
#pragma unroll
  for (int r = 0; r < NR; ++r) sv[r] = szs[r] = 0.0;
  // the first (up to) eight chunk partials and the z entries of a group,
  // loaded one group ahead so that the next group's loads are in flight
  // while this one is summed
  double xn[NR][8], zn[NR];
  auto load_head = [&](int g) {
    const int64_t c0 = a.group_chunk0[g], c1 = a.group_chunk0[g + 1];
    const double* z = th + 2 + 2 * K + static_cast<int64_t>(g) * (K + 1);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int ee = e + 32 * r;
#pragma unroll
      for (int u = 0; u < 8; ++u)
        xn[r][u] = (ee < K + 3 && c0 + u < c1) ? a.cpart[((c0 + u) * a.chains + c) * a.E + ee] : 0.0;
      zn[r] = ee <= K ? z[ee] : 0.0;
    }
  };
  if (g0 < g1) load_head(g0);
  for (int g = g0; g < g1; ++g) {
    const int64_t ch0 = a.group_chunk0[g], ch1 = a.group_chunk0[g + 1];
    double xc[NR][8], zc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      zc[r] = zn[r];
#pragma unroll
      for (int u = 0; u < 8; ++u) xc[r][u] = xn[r][u];
    }
    if (g + 1 < g1) load_head(g + 1);
    double zz = 0.0;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int ee = e + 32 * r;
      if (ee >= K + 3) continue;
      // chunks in order: the preloaded head, then eight loads in flight
      double v = 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (ch0 + u < ch1) v += xc[r][u];
      int64_t ch = ch0 + 8;
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
        const double zv = zc[r];
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
  for (int r = 0; r < NR; ++r) {
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

hipError_t launch_hier_x6(const hier_args& a, dim3 grid, hipStream_t s) {
  // no split copy of X, or no coefficient buffer
  if ((!a.XS || !a.wsplit) && a.n_chunks > 0) return hipErrorInvalidValue;
  const int64_t nc =
      static_cast<int64_t>((a.J + kCoefGroups - 1) / kCoefGroups) * a.chains * 32;
  if (nc > 0)
    hipLaunchKernelGGL(hier_coef_x6, dim3(static_cast<unsigned>((nc + 255) / 256)), dim3(256), 0,
                       s, a);
  const unsigned ncg8 = static_cast<unsigned>(((a.n_chunks + 3) / 4 + 7) / 8 * 8);
  grid = dim3(ncg8 * static_cast<unsigned>((a.chains + 31) / 32), 1);
  switch (a.K) {
#define SX_CASE(KK) \
  case KK: hipLaunchKernelGGL((hier_mfma_x6<KK>), grid, dim3(256), 0, s, a); break;
    SX_CASE(0) SX_CASE(1) SX_CASE(2) SX_CASE(3) SX_CASE(4) SX_CASE(5) SX_CASE(6) SX_CASE(7)
    SX_CASE(8) SX_CASE(9) SX_CASE(10) SX_CASE(11) SX_CASE(12) SX_CASE(13) SX_CASE(14)
    SX_CASE(15) SX_CASE(16) SX_CASE(17) SX_CASE(18) SX_CASE(19) SX_CASE(20) SX_CASE(21)
    SX_CASE(22) SX_CASE(23) SX_CASE(24) SX_CASE(25) SX_CASE(26) SX_CASE(27) SX_CASE(28)
    SX_CASE(29) SX_CASE(30) SX_CASE(31)
#undef SX_CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int NH>
hipError_t launch_hier_mfma(const hier_args& a, dim3 grid, hipStream_t s) {
  switch (a.K) {
#define SX_CASE(KK) \
  case KK: hipLaunchKernelGGL((hier_mfma_f32<KK, NH>), grid, dim3(256), 0, s, a); break;
    SX_CASE(0) SX_CASE(1) SX_CASE(2) SX_CASE(3) SX_CASE(4) SX_CASE(5) SX_CASE(6) SX_CASE(7)
    SX_CASE(8) SX_CASE(9) SX_CASE(10) SX_CASE(11) SX_CASE(12) SX_CASE(13) SX_CASE(14)
    SX_CASE(15) SX_CASE(16) SX_CASE(17) SX_CASE(18) SX_CASE(19) SX_CASE(20) SX_CASE(21)
    SX_CASE(22) SX_CASE(23) SX_CASE(24) SX_CASE(25) SX_CASE(26) SX_CASE(27) SX_CASE(28)
    SX_CASE(29) SX_CASE(30) SX_CASE(31) SX_CASE(32)
#undef SX_CASE
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
      switch (a.K) {
#define SX_CASE(KK) \
  case KK: hipLaunchKernelGGL((hier_pass_f64<KK>), grid, dim3(256), 0, s, a); break;
        SX_CASE(0) SX_CASE(1) SX_CASE(2) SX_CASE(3) SX_CASE(4) SX_CASE(5) SX_CASE(6) SX_CASE(7)
        SX_CASE(8) SX_CASE(9) SX_CASE(10) SX_CASE(11) SX_CASE(12) SX_CASE(13) SX_CASE(14)
        SX_CASE(15) SX_CASE(16) SX_CASE(17) SX_CASE(18) SX_CASE(19) SX_CASE(20) SX_CASE(21)
        SX_CASE(22) SX_CASE(23) SX_CASE(24) SX_CASE(25) SX_CASE(26) SX_CASE(27) SX_CASE(28)
        SX_CASE(29) SX_CASE(30) SX_CASE(31) SX_CASE(32)
#undef SX_CASE
        default: return hipErrorInvalidValue;
      }
    } else {
      const dim3 grid(static_cast<unsigned>((a.n_chunks + 3) / 4), static_cast<unsigned>((a.chains + 31) / 32));
      const hipError_t e =
          a.x6 && a.K <= 31 ? launch_hier_x6(a, grid, s) : launch_hier_mfma<1>(a, grid, s);
      if (e != hipSuccess) return e;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const dim3 ggrid((a.J + kGroupsPerBlock - 1) / kGroupsPerBlock, (a.chains + 7) / 8);
  if (a.K + 3 <= 32)
    hipLaunchKernelGGL(hier_groups<1>, ggrid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(hier_groups<2>, ggrid, dim3(256), 0, s, a);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(hier_final, dim3(a.chains), dim3(1024), 0, s, a);
  return hipGetLastError();
}

}  // namespace sx
