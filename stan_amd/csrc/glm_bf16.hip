// Batched GLM log density + gradient for many chains on MFMA (gfx950),
// X held as a BF16 hi+lo split with FP32 accumulation.
//
// For C chains with coefficients B (K x C) the forward pass is the dense
// contraction H = X B and the adjoint is G = X^T R with R the residual
// matrix: two GEMMs per evaluation, which is where MFMA pays (the single-
// chain pass is a matrix-vector product and stays on the HBM roofline).
// Each product of a split pair is formed as hi*hi + hi*lo + lo*hi (the
// lo*lo term is below the split's own representation error) with
// v_mfma_f32_32x32x16_bf16.
//
// The two contractions run over different indices of X (k forward, n in
// the adjoint) and an MFMA operand holds 8 consecutive inner-index values
// per lane, so X is stored twice, in HBM, as BF16 hi+lo:
//   X1[n][k]            k-packed rows      (forward operand)
//   X2[n/8][k][n % 8]   8-row-packed       (adjoint operand)
// Workgroup = 4 waves x 32 chains (glm_mfma32); a 1-D grid of row splits x
// chain blocks, mapped so that the blocks sharing a row range land on the
// same XCD and share X in its L2.  Per-block FP32 partials are summed
// across blocks in FP64 by a second launch that also adds priors and
// Jacobian (shared finalize).
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cmath>
#include <cstdint>

#include "glm_finalize.hpp"
#include "sx_internal.hpp"

namespace sx {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ __bf16 to_bf16(float x) { return static_cast<__bf16>(x); }

__device__ __forceinline__ void split(float x, __bf16* hi, __bf16* lo) {
  const __bf16 h = to_bf16(x);
  *hi = h;
  *lo = to_bf16(x - static_cast<float>(h));
}

// FP32 Bernoulli-logit row on the raw hardware transcendentals: term =
// -(max(-s eta, 0) + log(1 + t)), t = exp(-|s eta|), r = s inv_logit(-s eta).
// The exp2 argument is <= 0 and the log2 argument lies in [1, 2], where
// v_exp_f32 and v_log_f32 need none of the range handling of the library
// calls (an exp2 result below 2^-126 may flush to zero; it enters u = 1 + t
// and r = s t / u only at that size).  Absolute error of term ~1e-7,
// relative error of r a few ulp: below the BF16 hi+lo representation error
// of the products that feed eta.  s = 0 (a row past N) gives term = r = 0
// through the factor |s|, without a select.
__device__ __forceinline__ void logit_row_fast(float eta, float s, float* term, float* r) {
  constexpr float kLog2e = 1.44269504088896341f, kLn2 = 0.693147180559945309f;
  const float se = s * eta;
  const float t = __builtin_amdgcn_exp2f(-fabsf(se) * kLog2e);
  const float u = 1.0f + t;
  const float inv = __builtin_amdgcn_rcpf(u);
  *term = -fabsf(s) * fmaf(__builtin_amdgcn_logf(u), kLn2, fmaxf(-se, 0.0f));
  *r = s * (se >= 0.0f ? t * inv : inv);
}

// Splits row-major FP64 X into the two BF16 hi+lo layouts (zero padded).
__global__ void split_x(const double* __restrict__ Xrm, __bf16* X1h, __bf16* X1l, __bf16* X2h,
                        __bf16* X2l, int64_t N, int64_t K, int64_t Kp, int64_t total) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int64_t n = idx / Kp;
  const int64_t k = idx - n * Kp;
  const float x = (n < N && k < K) ? static_cast<float>(Xrm[n * K + k]) : 0.0f;
  __bf16 h, l;
  split(x, &h, &l);
  X1h[idx] = h;
  X1l[idx] = l;
  const int64_t j = ((n >> 3) * Kp + k) * 8 + (n & 7);
  X2h[j] = h;
  X2l[j] = l;
}

// theta (C x D, FP64) -> B hi/lo (Cp x Kp), alpha (Cp)
__global__ void prep_chains(const double* __restrict__ theta, int64_t D, int C, int K, int Kp,
                            __bf16* Bh, __bf16* Bl, float* alpha) {
  const int c = blockIdx.x;
  for (int k = threadIdx.x; k < Kp; k += blockDim.x) {
    const float b = (c < C && k < K) ? static_cast<float>(theta[c * D + 1 + k]) : 0.0f;
    __bf16 h, l;
    split(b, &h, &l);
    Bh[static_cast<int64_t>(c) * Kp + k] = h;
    Bl[static_cast<int64_t>(c) * Kp + k] = l;
  }
  if (threadIdx.x == 0) alpha[c] = c < C ? static_cast<float>(theta[c * D]) : 0.0f;
}

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));  // a 16-byte staging block
typedef __attribute__((address_space(3))) void* lds_ptr;  // LDS pointer of global_load_lds

// K <= 256 (KQ = Kp / 32 <= 8): workgroup = NW waves x 32 chains (NW = 4
// for 128-chain tiles, 2 for batches of at most 64 chains).  Each 32-row tile of X (both layouts, hi and lo) is staged once
// per workgroup in LDS, straight from global memory (global_load_lds), and
// read by all four waves, so X crosses L2 once per 128 chains.  A software
// pipeline over tiles: while the link and adjoint of tile t run, the
// forward MFMAs of tile t + 1 (independent of them) interleave with the
// link's VALU work, X1 of tile t + 2 and X2 of tile t + 1 stream in; two
// buffers per layout suffice.  v_mfma_f32_32x32x16_bf16 operands (lane l,
// j < 8): A[l % 32][8 (l / 32) + j], B[8 (l / 32) + j][l % 32]; result
// d[j] = D[8 (j / 4) + 4 (l / 32) + j % 4][l % 32].
//   forward  H (32 rows x 32 chains) = X1 tile (A, from LDS) . B (registers)
//            lane l holds chain l % 32 and rows 8 (j/4) + 4 (l/32) + j%4
//   adjoint  G^T (32 chains x 32 k) += R^T . X: the contraction runs over
//            rows in the order the forward left them in the lane (step st
//            uses the lane's own r[8st .. 8st+7]), and the X operand takes
//            the same rows from the 8-row-packed tile: rows 16st + 4 (l/32)
//            + 0..3 of groups 2st and 2st+1 (two 8-byte LDS reads).
template <int KQ, int FAM, int NW>
__global__ __launch_bounds__(64 * NW) void glm_mfma32(mx_args a) {
  static_assert(KQ >= 1 && KQ <= 8, "glm_mfma32 handles 32 <= Kp <= 256");
  static_assert(NW == 2 || NW == 4, "2 or 4 waves per workgroup");
  constexpr int Kp = 32 * KQ;
  constexpr int S16 = 2 * KQ;          // forward k-steps of 16
  constexpr int CPR = Kp / 8;          // 16-byte chunks per X1 row
  // XOR swizzle of X1 chunks: the largest power of two (<= 16) dividing
  // CPR, so c ^ (r & SWZ) stays inside the row
  constexpr int SWZ = (CPR % 16 == 0 ? 16 : CPR % 8 == 0 ? 8 : 4) - 1;
  constexpr int TE = 32 * Kp;          // elements of one tile array
  constexpr int BPA = Kp / 16;         // 1 KB blocks per tile array
  __shared__ __attribute__((aligned(16))) __bf16 x1[2][2][TE];  // [buffer][hi, lo]
  __shared__ __attribute__((aligned(16))) __bf16 x2[2][2][TE];
  __shared__ float ys[2][32];  // y of the tile's rows (+-1 for the logit)
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 31, h = lane >> 5;
  // XCD-aware 1-D grid: the splits of one chain block are 8 block ids
  // apart, so blocks sharing a row range run on the same XCD and L2
  const int nCB = a.Cp / (32 * NW);
  const int bid = blockIdx.x;
  const int xcd = bid & 7, widx = bid >> 3;
  const int spl = (widx / nCB) * 8 + xcd;
  const int cblk = widx % nCB;
  const int cbase = cblk * 32 * NW + 32 * w;  // first chain of this wave
  const int chain = cbase + li;
  bf16x8 bh[S16], bl[S16];
#pragma unroll
  for (int s = 0; s < S16; ++s) {
    const int64_t off = static_cast<int64_t>(chain) * Kp + 16 * s + 8 * h;
    bh[s] = *reinterpret_cast<const bf16x8*>(a.Bh + off);
    bl[s] = *reinterpret_cast<const bf16x8*>(a.Bl + off);
  }
  const float alpha = a.alpha[chain];
  f32x16 g[KQ];
#pragma unroll
  for (int i = 0; i < KQ; ++i)
#pragma unroll
    for (int j = 0; j < 16; ++j) g[i][j] = 0.f;
  double lp = 0.0, sr = 0.0;
  float nf = 0.f;  // sum of 0 * eta over the lane's rows

  const int64_t t0 = static_cast<int64_t>(spl) * a.tiles_per_split;
  const int64_t t1 = t0 + a.tiles_per_split < a.n_tiles ? t0 + a.tiles_per_split : a.n_tiles;
  const __bf16* X1h = reinterpret_cast<const __bf16*>(a.X1h);
  const __bf16* X1l = reinterpret_cast<const __bf16*>(a.X1l);
  const __bf16* X2h = reinterpret_cast<const __bf16*>(a.X2h);
  const __bf16* X2l = reinterpret_cast<const __bf16*>(a.X2l);
  double yraw = 0.0;  // y of a staged row (tid < 32): converted when committed
  int yraw_i = 0;
  bool yvalid = false;
  float ystage = 0.f;  // (direct staging: converted at once)
  // Staging.  Up to K = 192 (KQ <= 6) through registers: a tile's X1 (X2,
  // with its y) is loaded with ordinary 16-byte global loads at the top of
  // an iteration and written to its LDS buffer at the end, so the loads
  // have the iteration's compute to land and the only wait for them is at
  // those writes.  Above, where the loaded blocks would not fit beside B
  // hi/lo and the gradient accumulators, straight from global memory to LDS
  // (global_load_lds), waited for before the barrier (the compiler also
  // waits for them before the next LDS read, so these do not overlap the
  // tile's compute).  One layout's tile is 2 BPA blocks of 1 KB (hi, then
  // lo); wave w moves blocks w, w + NW, ... (4 KQ / NW of them), lane l 16
  // bytes at block offset 16 l.  X1 rows are stored with chunk c of row r at
  // c ^ (r & SWZ), so the forward's row-per-lane reads spread over the
  // banks; X2 is copied as is.  y is loaded raw and converted when
  // committed, so that no early use of its register forces a wait for it
  // (and for every load issued before it).
  constexpr bool REG = KQ <= 6;
  constexpr int NBW = 4 * KQ / NW;     // staging blocks per wave and layout
  u32x4 r1[REG ? NBW : 1], r2[REG ? NBW : 1];
#define SX_LOAD_X1(T, BUF)                                                       \
  {                                                                              \
    const int64_t base_ = (T) * 32 * Kp;                                         \
    _Pragma("clang loop unroll(full)") for (int i = 0; i < NBW; ++i) {           \
      const int b_ = w + NW * i;                                                 \
      const int p_ = (b_ % BPA) * 64 + lane;                                     \
      const int r_ = p_ / CPR;                                                   \
      const int src_ = r_ * CPR + ((p_ % CPR) ^ (r_ & SWZ));                     \
      const __bf16* g_ = (b_ < BPA ? X1h : X1l) + base_ + 8 * src_;              \
      if constexpr (REG)                                                         \
        r1[i] = *reinterpret_cast<const u32x4*>(g_);                             \
      else                                                                       \
        __builtin_amdgcn_global_load_lds(g_,                                     \
            (lds_ptr)(&x1[BUF][b_ / BPA][(b_ % BPA) * 512]), 16, 0, 0);          \
    }                                                                            \
  }
#define SX_LOAD_X2(T, BUF)                                                       \
  {                                                                              \
    const int64_t base_ = (T) * 32 * Kp;                                         \
    _Pragma("clang loop unroll(full)") for (int i = 0; i < NBW; ++i) {           \
      const int b_ = w + NW * i;                                                 \
      const int p_ = (b_ % BPA) * 64 + lane;                                     \
      const __bf16* g_ = (b_ < BPA ? X2h : X2l) + base_ + 8 * p_;                \
      if constexpr (REG)                                                         \
        r2[i] = *reinterpret_cast<const u32x4*>(g_);                             \
      else                                                                       \
        __builtin_amdgcn_global_load_lds(g_,                                     \
            (lds_ptr)(&x2[BUF][b_ / BPA][(b_ % BPA) * 512]), 16, 0, 0);          \
    }                                                                            \
    if (tid < 32) {                                                              \
      const int64_t n_ = (T) * 32 + tid;                                         \
      yvalid = n_ < a.N;                                                         \
      const int64_t m_ = yvalid ? n_ : 0;                                        \
      if (FAM == SX_LOGIT)                                                       \
        yraw_i = static_cast<const int8_t*>(a.y)[m_];                            \
      else                                                                       \
        yraw = static_cast<const double*>(a.y)[m_];                              \
      if constexpr (!REG)                                                        \
        ystage = !yvalid ? 0.f : FAM == SX_LOGIT ? (yraw_i ? 1.0f : -1.0f)      \
                                                 : static_cast<float>(yraw);     \
    }                                                                            \
  }
#define SX_COMMIT_X1(BUF)                                                        \
  {                                                                              \
    if constexpr (REG) {                                                         \
      _Pragma("clang loop unroll(full)") for (int i = 0; i < NBW; ++i) {         \
        const int b_ = w + NW * i;                                               \
        *reinterpret_cast<u32x4*>(&x1[BUF][b_ / BPA][(b_ % BPA) * 512 + 8 * lane]) = r1[i]; \
      }                                                                          \
    }                                                                            \
  }
  // X2 and the tile's y; the direct loads of both layouts are waited for
#define SX_COMMIT_X2(BUF)                                                        \
  {                                                                              \
    if constexpr (REG) {                                                         \
      _Pragma("clang loop unroll(full)") for (int i = 0; i < NBW; ++i) {         \
        const int b_ = w + NW * i;                                               \
        *reinterpret_cast<u32x4*>(&x2[BUF][b_ / BPA][(b_ % BPA) * 512 + 8 * lane]) = r2[i]; \
      }                                                                          \
    } else {                                                                     \
      __builtin_amdgcn_s_waitcnt(0);                                             \
    }                                                                            \
    if (tid < 32) {                                                              \
      float v_ = ystage;                                                         \
      if constexpr (REG)                                                         \
        v_ = !yvalid ? 0.f : FAM == SX_LOGIT ? (yraw_i ? 1.0f : -1.0f)          \
                                             : static_cast<float>(yraw);         \
      ys[BUF][tid] = v_;                                                         \
    }                                                                            \
  }
#define SX_FORWARD(BUF, ACC)                                                     \
  {                                                                              \
    _Pragma("unroll") for (int j = 0; j < 16; ++j) ACC[j] = 0.f;                 \
    _Pragma("unroll") for (int s = 0; s < S16; ++s) {                            \
      const int off = 8 * (li * CPR + ((2 * s + h) ^ (li & SWZ)));               \
      const bf16x8 xh = *reinterpret_cast<const bf16x8*>(&x1[BUF][0][off]);      \
      const bf16x8 xl = *reinterpret_cast<const bf16x8*>(&x1[BUF][1][off]);      \
      ACC = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, bh[s], ACC, 0, 0, 0);    \
      ACC = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, bl[s], ACC, 0, 0, 0);    \
      ACC = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xl, bh[s], ACC, 0, 0, 0);    \
    }                                                                            \
  }
  f32x16 hcur, hnext;
#pragma unroll
  for (int j = 0; j < 16; ++j) hcur[j] = hnext[j] = 0.f;
  // prologue: X1 and X2 of tile t0, X1 of tile t0 + 1; forward of tile t0
  if (t0 < t1) {
    SX_LOAD_X1(t0, 0);
    SX_COMMIT_X1(0);
    if (t0 + 1 < t1) {
      SX_LOAD_X1(t0 + 1, 1);
      SX_COMMIT_X1(1);
    }
    SX_LOAD_X2(t0, 0);
    SX_COMMIT_X2(0);
  }
  __syncthreads();
  if (t0 < t1) SX_FORWARD(0, hcur);
  for (int64_t t = t0; t < t1; ++t) {
    const int cur = static_cast<int>((t - t0) & 1), nxt = cur ^ 1;
    // x1[cur] held X1 of tile t (its forward ran an iteration ago) and
    // x2[nxt] X2 of tile t - 1: both free since the last barrier.  The
    // forward of the tile after the last reads a stale buffer; its result
    // is discarded.
    if (t + 2 < t1) SX_LOAD_X1(t + 2, cur);
    if (t + 1 < t1) SX_LOAD_X2(t + 1, nxt);
    SX_FORWARD(nxt, hnext);
    // link and adjoint of tile t in two halves of the lane's 16 rows: half
    // st holds elements j = 8 st .. 8 st + 7 (rows 8 (j/4) + 4 h + j%4),
    // the rows the adjoint's step st contracts.  Rows past N are masked.
    // The tile's sums are FP32 and are added to the FP64 totals per tile.
    const int64_t n0 = t * 32;
    float lpt = 0.f, srt = 0.f;
#pragma unroll
    for (int st = 0; st < 2; ++st) {
      bf16x8 rh, rl;
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const int j = 8 * st + jj;
        const int row = 8 * (j >> 2) + 4 * h + (j & 3);
        const float yv = ys[cur][row];
        const float eta = hcur[j] + alpha;
        float term, rr;
        if (FAM == SX_LOGIT) {
          // rows past N carry sign 0: term = r = 0
          logit_row_fast(eta, yv, &term, &rr);
        } else {
          rr = n0 + row < a.N ? yv - eta : 0.f;
          term = rr * rr;
        }
        // NaN once any eta is infinite or NaN; a row past N has eta =
        // alpha, non-finite only if every row's is
        nf = fmaf(eta, 0.f, nf);
        lpt += term;
        srt += rr;
        __bf16 hh, ll;
        split(rr, &hh, &ll);
        rh[jj] = hh;
        rl[jj] = ll;
      }
#pragma unroll
      for (int kb = 0; kb < KQ; ++kb) {
        const int k = 32 * kb + li;
        const int o0 = ((2 * st) * Kp + k) * 8 + 4 * h;
        const int o1 = ((2 * st + 1) * Kp + k) * 8 + 4 * h;
        const bf16x4 ah0 = *reinterpret_cast<const bf16x4*>(&x2[cur][0][o0]);
        const bf16x4 ah1 = *reinterpret_cast<const bf16x4*>(&x2[cur][0][o1]);
        const bf16x4 al0 = *reinterpret_cast<const bf16x4*>(&x2[cur][1][o0]);
        const bf16x4 al1 = *reinterpret_cast<const bf16x4*>(&x2[cur][1][o1]);
        bf16x8 xh, xl;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          xh[j] = ah0[j];
          xh[4 + j] = ah1[j];
          xl[j] = al0[j];
          xl[4 + j] = al1[j];
        }
        g[kb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh, xh, g[kb], 0, 0, 0);
        g[kb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh, xl, g[kb], 0, 0, 0);
        g[kb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rl, xh, g[kb], 0, 0, 0);
      }
    }
    lp += static_cast<double>(lpt);
    sr += static_cast<double>(srt);

    if (t + 2 < t1) SX_COMMIT_X1(cur);
    if (t + 1 < t1) SX_COMMIT_X2(nxt);
    __syncthreads();
    hcur = hnext;
  }
#undef SX_FORWARD
#undef SX_LOAD_X1
#undef SX_LOAD_X2
#undef SX_COMMIT_X1
#undef SX_COMMIT_X2
  // g[kb][j] = G^T[chain cbase + 8 (j/4) + 4 h + j%4][k = 32 kb + li]
  const int64_t pbase = static_cast<int64_t>(spl) * a.Cp;
#pragma unroll
  for (int kb = 0; kb < KQ; ++kb)
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int ci = cbase + 8 * (j >> 2) + 4 * h + (j & 3);
      a.gpart[(pbase + ci) * Kp + 32 * kb + li] = g[kb][j];
    }
  // row terms of this lane's chain: combine the two halves of the rows
  lp += __shfl_xor(lp, 32, 64);
  sr += __shfl_xor(sr, 32, 64);
  nf += __shfl_xor(nf, 32, 64);
  const double bad = nf == nf ? 0.0 : 1.0;
  if (h == 0) {
    double* sp = a.spart + (pbase + chain) * 3;
    sp[0] = lp;
    sp[1] = sr;
    sp[2] = bad;
  }
}

// 256 < K <= 512 (KQ = Kp / 32 in {10, 12, 14, 16}): B hi/lo of 32 chains
// over the full K fill 256 registers, so the gradient accumulators cannot
// stay beside them; the evaluation is split in two launches.  This one is
// glm_mfma32's forward and link (same workgroup, grid, X1 staging and
// software pipeline): it writes the residual r of every row and chain as
// the BF16 hi/lo MFMA operands the adjoint will use, lane by lane (tile t,
// 32-chain group g, half st: the lane's bf16x8 at
// R[((t * Cp / 32 + g) * 2 + st) * 64 + lane]), and the row terms per
// split.  glm_mfma32_adj then forms G^T = R^T X per block of 256 k.
template <int KQ, int FAM, int NW>
__global__ __launch_bounds__(64 * NW) void glm_mfma32_fwd(mx_args a) {
  static_assert(KQ > 8 && KQ <= 16 && KQ % 2 == 0, "glm_mfma32_fwd handles 256 < Kp <= 512, Kp % 64 == 0");
  static_assert(NW == 2 || NW == 4, "2 or 4 waves per workgroup");
  constexpr int Kp = 32 * KQ;
  constexpr int S16 = 2 * KQ;
  constexpr int CPR = Kp / 8;
  constexpr int SWZ = (CPR % 16 == 0 ? 16 : CPR % 8 == 0 ? 8 : 4) - 1;
  constexpr int TE = 32 * Kp;
  constexpr int BPA = Kp / 16;
  __shared__ __attribute__((aligned(16))) __bf16 x1[2][2][TE];  // [buffer][hi, lo]
  __shared__ float ys[2][32];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 31, h = lane >> 5;
  const int nCB = a.Cp / (32 * NW);
  const int bid = blockIdx.x;
  const int xcd = bid & 7, widx = bid >> 3;
  const int spl = (widx / nCB) * 8 + xcd;
  const int cblk = widx % nCB;
  const int cbase = cblk * 32 * NW + 32 * w;
  const int chain = cbase + li;
  const int64_t g32 = cbase / 32, nG = a.Cp / 32;
  bf16x8 bh[S16], bl[S16];
#pragma unroll
  for (int s = 0; s < S16; ++s) {
    const int64_t off = static_cast<int64_t>(chain) * Kp + 16 * s + 8 * h;
    bh[s] = *reinterpret_cast<const bf16x8*>(a.Bh + off);
    bl[s] = *reinterpret_cast<const bf16x8*>(a.Bl + off);
  }
  const float alpha = a.alpha[chain];
  double lp = 0.0, sr = 0.0;
  float nf = 0.f;  // sum of 0 * eta over the lane's rows
  const int64_t t0 = static_cast<int64_t>(spl) * a.tiles_per_split;
  const int64_t t1 = t0 + a.tiles_per_split < a.n_tiles ? t0 + a.tiles_per_split : a.n_tiles;
  const __bf16* X1h = reinterpret_cast<const __bf16*>(a.X1h);
  const __bf16* X1l = reinterpret_cast<const __bf16*>(a.X1l);
  bf16x8* Rh = reinterpret_cast<bf16x8*>(a.Rh);
  bf16x8* Rl = reinterpret_cast<bf16x8*>(a.Rl);
  float ystage = 0.f;
#define SX_STAGE_X1(T, BUF)                                                      \
  {                                                                              \
    const int64_t base_ = (T) * 32 * Kp;                                         \
    _Pragma("unroll") for (int i = 0; i < 4 * KQ / NW; ++i) {                    \
      const int b_ = w + NW * i;                                                 \
      const int p_ = (b_ % BPA) * 64 + lane;                                     \
      const int r_ = p_ / CPR;                                                   \
      const int src_ = r_ * CPR + ((p_ % CPR) ^ (r_ & SWZ));                     \
      __builtin_amdgcn_global_load_lds((b_ < BPA ? X1h : X1l) + base_ + 8 * src_, \
          (lds_ptr)(&x1[BUF][b_ / BPA][(b_ % BPA) * 512]), 16, 0, 0);            \
    }                                                                            \
  }
#define SX_STAGE_Y(T)                                                            \
  {                                                                              \
    if (tid < 32) {                                                              \
      const int64_t n_ = (T) * 32 + tid;                                         \
      float v_ = 0.f;                                                            \
      if (n_ < a.N) {                                                            \
        if (FAM == SX_LOGIT)                                                     \
          v_ = static_cast<const int8_t*>(a.y)[n_] ? 1.0f : -1.0f;               \
        else                                                                     \
          v_ = static_cast<float>(static_cast<const double*>(a.y)[n_]);          \
      }                                                                          \
      ystage = v_;                                                               \
    }                                                                            \
  }
#define SX_STAGE_DONE(YBUF, WITH_Y)                                              \
  {                                                                              \
    __builtin_amdgcn_s_waitcnt(0);                                               \
    if ((WITH_Y) && tid < 32) ys[YBUF][tid] = ystage;                            \
  }
#define SX_FORWARD(BUF, ACC)                                                     \
  {                                                                              \
    _Pragma("unroll") for (int j = 0; j < 16; ++j) ACC[j] = 0.f;                 \
    _Pragma("unroll") for (int s = 0; s < S16; ++s) {                            \
      const int off = 8 * (li * CPR + ((2 * s + h) ^ (li & SWZ)));               \
      const bf16x8 xh = *reinterpret_cast<const bf16x8*>(&x1[BUF][0][off]);      \
      const bf16x8 xl = *reinterpret_cast<const bf16x8*>(&x1[BUF][1][off]);      \
      ACC = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, bh[s], ACC, 0, 0, 0);    \
      ACC = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xh, bl[s], ACC, 0, 0, 0);    \
      ACC = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xl, bh[s], ACC, 0, 0, 0);    \
    }                                                                            \
  }
  f32x16 hcur, hnext;
#pragma unroll
  for (int j = 0; j < 16; ++j) hcur[j] = hnext[j] = 0.f;
  if (t0 < t1) {
    SX_STAGE_X1(t0, 0);
    if (t0 + 1 < t1) SX_STAGE_X1(t0 + 1, 1);
    SX_STAGE_Y(t0);
    SX_STAGE_DONE(0, true);
  }
  __syncthreads();
  if (t0 < t1) SX_FORWARD(0, hcur);
  for (int64_t t = t0; t < t1; ++t) {
    const int cur = static_cast<int>((t - t0) & 1), nxt = cur ^ 1;
    if (t + 2 < t1) SX_STAGE_X1(t + 2, cur);
    if (t + 1 < t1) SX_STAGE_Y(t + 1);
    SX_FORWARD(nxt, hnext);
    const int64_t n0 = t * 32;
    float lpt = 0.f, srt = 0.f;
#pragma unroll
    for (int st = 0; st < 2; ++st) {
      bf16x8 rh, rl;
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const int j = 8 * st + jj;
        const int row = 8 * (j >> 2) + 4 * h + (j & 3);
        const float yv = ys[cur][row];
        const float eta = hcur[j] + alpha;
        float term, rr;
        if (FAM == SX_LOGIT) {
          // rows past N carry sign 0: term = r = 0
          logit_row_fast(eta, yv, &term, &rr);
        } else {
          rr = n0 + row < a.N ? yv - eta : 0.f;
          term = rr * rr;
        }
        // NaN once any eta is infinite or NaN; a row past N has eta =
        // alpha, non-finite only if every row's is
        nf = fmaf(eta, 0.f, nf);
        lpt += term;
        srt += rr;
        __bf16 hh, ll;
        split(rr, &hh, &ll);
        rh[jj] = hh;
        rl[jj] = ll;
      }
      const int64_t ri = ((t * nG + g32) * 2 + st) * 64 + lane;
      Rh[ri] = rh;
      Rl[ri] = rl;
    }
    lp += static_cast<double>(lpt);
    sr += static_cast<double>(srt);

    SX_STAGE_DONE(nxt, t + 1 < t1);
    __syncthreads();
    hcur = hnext;
  }
#undef SX_FORWARD
#undef SX_STAGE_X1
#undef SX_STAGE_Y
#undef SX_STAGE_DONE
  const int64_t pbase = static_cast<int64_t>(spl) * a.Cp;
  lp += __shfl_xor(lp, 32, 64);
  sr += __shfl_xor(sr, 32, 64);
  nf += __shfl_xor(nf, 32, 64);
  const double bad = nf == nf ? 0.0 : 1.0;
  if (h == 0) {
    double* sp = a.spart + (pbase + chain) * 3;
    sp[0] = lp;
    sp[1] = sr;
    sp[2] = bad;
  }
}

// The adjoint of a 256 < K <= 512 evaluation for one block of k (KQA x 32
// columns from column 32 kq0): G^T (32 chains x 32 k) += R^T X per tile,
// as in glm_mfma32 (step st takes the lane's residuals of rows 8 (j/4) +
// 4 h + j%4, j = 8 st .. 8 st + 7, and the same rows of the 8-row-packed
// tile), with R read from the forward launch's output and the tile's
// column block of X2 staged through LDS one tile ahead.
template <int KQA, int NW>
__global__ __launch_bounds__(64 * NW) void glm_mfma32_adj(mx_args a, int kq0) {
  static_assert(KQA >= 1 && KQA <= 8, "blocks of at most 256 columns");
  static_assert((4 * KQA) % NW == 0, "staging blocks per wave");
  constexpr int KR = 32 * KQA;         // columns of the block
  constexpr int TE = 32 * KR;          // elements of one tile array
  constexpr int NB = 2 * KQA;          // 1 KB staging blocks per tile array
  __shared__ __attribute__((aligned(16))) __bf16 x2[2][2][TE];  // [buffer][hi, lo]
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 31, h = lane >> 5;
  const int nCB = a.Cp / (32 * NW);
  const int bid = blockIdx.x;
  const int xcd = bid & 7, widx = bid >> 3;
  const int spl = (widx / nCB) * 8 + xcd;
  const int cblk = widx % nCB;
  const int cbase = cblk * 32 * NW + 32 * w;
  const int64_t g32 = cbase / 32, nG = a.Cp / 32;
  const int Kp = a.Kp;
  const int k0 = 32 * kq0;
  f32x16 g[KQA];
#pragma unroll
  for (int i = 0; i < KQA; ++i)
#pragma unroll
    for (int j = 0; j < 16; ++j) g[i][j] = 0.f;
  const int64_t t0 = static_cast<int64_t>(spl) * a.tiles_per_split;
  const int64_t t1 = t0 + a.tiles_per_split < a.n_tiles ? t0 + a.tiles_per_split : a.n_tiles;
  const __bf16* X2h = reinterpret_cast<const __bf16*>(a.X2h);
  const __bf16* X2l = reinterpret_cast<const __bf16*>(a.X2l);
  const bf16x8* Rh = reinterpret_cast<const bf16x8*>(a.Rh);
  const bf16x8* Rl = reinterpret_cast<const bf16x8*>(a.Rl);
  // staging: chunk p (16 bytes: the 8 rows of one column) of the tile's
  // column block is group p / KR, column k0 + p % KR; wave w moves blocks
  // w, w + NW, ... of 64 chunks per array (hi, then lo)
#define SX_STAGE_X2B(T, BUF)                                                     \
  {                                                                              \
    _Pragma("unroll") for (int i = 0; i < 2 * NB / NW; ++i) {                    \
      const int b_ = w + NW * i;                                                 \
      const int bb_ = b_ % NB;                                                   \
      const int p_ = bb_ * 64 + lane;                                            \
      const int64_t src_ = ((4 * (T) + p_ / KR) * Kp + k0 + p_ % KR) * 8;        \
      __builtin_amdgcn_global_load_lds((b_ < NB ? X2h : X2l) + src_,             \
          (lds_ptr)(&x2[BUF][b_ / NB][bb_ * 512]), 16, 0, 0);                    \
    }                                                                            \
  }
  bf16x8 rh[2], rl[2], nh[2], nl[2];
  if (t0 < t1) {
    SX_STAGE_X2B(t0, 0);
#pragma unroll
    for (int st = 0; st < 2; ++st) {
      const int64_t ri = ((t0 * nG + g32) * 2 + st) * 64 + lane;
      rh[st] = Rh[ri];
      rl[st] = Rl[ri];
    }
    __builtin_amdgcn_s_waitcnt(0);
  }
  __syncthreads();
  for (int64_t t = t0; t < t1; ++t) {
    const int cur = static_cast<int>((t - t0) & 1), nxt = cur ^ 1;
    // x2[nxt] held tile t - 1, read before the last barrier
    if (t + 1 < t1) {
      SX_STAGE_X2B(t + 1, nxt);
#pragma unroll
      for (int st = 0; st < 2; ++st) {
        const int64_t ri = (((t + 1) * nG + g32) * 2 + st) * 64 + lane;
        nh[st] = Rh[ri];
        nl[st] = Rl[ri];
      }
    }
#pragma unroll
    for (int st = 0; st < 2; ++st) {
#pragma unroll
      for (int kb = 0; kb < KQA; ++kb) {
        const int k = 32 * kb + li;
        const int o0 = ((2 * st) * KR + k) * 8 + 4 * h;
        const int o1 = ((2 * st + 1) * KR + k) * 8 + 4 * h;
        const bf16x4 ah0 = *reinterpret_cast<const bf16x4*>(&x2[cur][0][o0]);
        const bf16x4 ah1 = *reinterpret_cast<const bf16x4*>(&x2[cur][0][o1]);
        const bf16x4 al0 = *reinterpret_cast<const bf16x4*>(&x2[cur][1][o0]);
        const bf16x4 al1 = *reinterpret_cast<const bf16x4*>(&x2[cur][1][o1]);
        bf16x8 xh, xl;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          xh[j] = ah0[j];
          xh[4 + j] = ah1[j];
          xl[j] = al0[j];
          xl[4 + j] = al1[j];
        }
        g[kb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh[st], xh, g[kb], 0, 0, 0);
        g[kb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh[st], xl, g[kb], 0, 0, 0);
        g[kb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rl[st], xh, g[kb], 0, 0, 0);
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
#pragma unroll
    for (int st = 0; st < 2; ++st) {
      rh[st] = nh[st];
      rl[st] = nl[st];
    }
  }
#undef SX_STAGE_X2B
  const int64_t pbase = static_cast<int64_t>(spl) * a.Cp;
#pragma unroll
  for (int kb = 0; kb < KQA; ++kb)
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int ci = cbase + 8 * (j >> 2) + 4 * h + (j & 3);
      a.gpart[(pbase + ci) * Kp + k0 + 32 * kb + li] = g[kb][j];
    }
}

// Per chain: FP64 sums of the split partials, then priors and output.  A
// split's gradient partials are a contiguous row of Kp floats; the block's
// 1024 threads cover the rows as Kp/4 float4 columns x 4096/Kp slices of
// splits, every thread active and eight 16-byte loads in flight.  Each
// thread sums its splits in ascending order, the slices are combined in
// slice order: a fixed order that depends on N, K and the split count only.
template <int FAM>
__global__ __launch_bounds__(1024) void mx_reduce(mx_args a, pass_args pa) {
  extern __shared__ double st[];
  __shared__ double red[4096];  // [slice][k], Kp x slices = 4096
  __shared__ double rs[3][64];
  const int c = blockIdx.x;
  pa.theta += static_cast<int64_t>(c) * pa.theta_stride;
  pa.out += static_cast<int64_t>(c) * pa.out_stride;
  const int K = pa.K, Kp = a.Kp;
  const int q = Kp / 4;               // float4 columns of a row
  const int nsl = 1024 / q;           // slices of splits (q * nsl <= 1024)
  const int tid = threadIdx.x;
  const int kq = tid % q, sl = tid / q;
  const int64_t gstride = static_cast<int64_t>(a.Cp) * Kp;
  const float* gp = a.gpart + static_cast<int64_t>(c) * Kp + 4 * kq;
  double v[4] = {0, 0, 0, 0};
  // threads past q * nsl (Kp not dividing 4096) take no splits
  int si = sl < nsl ? sl : a.splits;
  for (; si + 7 * nsl < a.splits; si += 8 * nsl) {
    float4 x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      x[u] = *reinterpret_cast<const float4*>(gp + static_cast<int64_t>(si + u * nsl) * gstride);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      v[0] += static_cast<double>(x[u].x);
      v[1] += static_cast<double>(x[u].y);
      v[2] += static_cast<double>(x[u].z);
      v[3] += static_cast<double>(x[u].w);
    }
  }
  for (; si < a.splits; si += nsl) {
    const float4 x = *reinterpret_cast<const float4*>(gp + static_cast<int64_t>(si) * gstride);
    v[0] += static_cast<double>(x.x);
    v[1] += static_cast<double>(x.y);
    v[2] += static_cast<double>(x.z);
    v[3] += static_cast<double>(x.w);
  }
  if (sl < nsl)
#pragma unroll
    for (int i = 0; i < 4; ++i) red[sl * Kp + 4 * kq + i] = v[i];
  // row terms (lp, sum r, bad): entry e on threads 64 e .. 64 e + 63, each
  // over splits l, l + 64, ... (eight loads in flight), then the 64 in order
  if (tid < 192) {
    const int e = tid >> 6, l = tid & 63;
    const double* sp = a.spart + static_cast<int64_t>(c) * 3 + e;
    const int64_t sstride = 3 * static_cast<int64_t>(a.Cp);
    double t = 0.0;
    int s2 = l;
    for (; s2 + 7 * 64 < a.splits; s2 += 8 * 64) {
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) x[u] = sp[static_cast<int64_t>(s2 + 64 * u) * sstride];
#pragma unroll
      for (int u = 0; u < 8; ++u) t += x[u];
    }
    for (; s2 < a.splits; s2 += 64) t += sp[static_cast<int64_t>(s2) * sstride];
    rs[e][l] = t;
  }
  __syncthreads();
  if (tid < K) {
    double t = 0.0;
    for (int s2 = 0; s2 < nsl; ++s2) t += red[s2 * Kp + tid];
    st[2 + tid] = t;  // finalize layout: st[0] lp, st[1] sum r, st[2 + k], st[K + 2] bad
  } else if (tid >= 1021) {
    const int e = tid - 1021;
    double t = 0.0;
    for (int l = 0; l < 64; ++l) t += rs[e][l];
    st[e == 0 ? 0 : (e == 1 ? 1 : K + 2)] = t;
  }
  __syncthreads();
  if (tid < 64) finalize<FAM>(pa, st);
}

template <int KQ, int NW>
hipError_t launch_kq32_nw(int family, const mx_args& a, hipStream_t s) {
  if (a.splits % 8 != 0 || a.Cp % (32 * NW) != 0) return hipErrorInvalidValue;
  dim3 grid(a.splits * (a.Cp / (32 * NW)));
  if (family == SX_LOGIT)
    hipLaunchKernelGGL((glm_mfma32<KQ, SX_LOGIT, NW>), grid, dim3(64 * NW), 0, s, a);
  else
    hipLaunchKernelGGL((glm_mfma32<KQ, SX_NORMAL, NW>), grid, dim3(64 * NW), 0, s, a);
  return hipGetLastError();
}

// chain tile: a batch of at most 64 chains (Cp = 64) gets two waves per
// workgroup, larger batches four.  (One-wave workgroups of 32 chains were
// measured slower than two-wave ones of 64: 0.274 vs 0.257 ms at N = 1e6,
// K = 100.)
template <int KQ>
hipError_t launch_kq32(int family, const mx_args& a, hipStream_t s) {
  if (a.Cp == 64) return launch_kq32_nw<KQ, 2>(family, a, s);
  return launch_kq32_nw<KQ, 4>(family, a, s);
}

// 256 < Kp <= 512: forward + link, then the adjoint per block of 256 k
template <int KQ, int NW>
hipError_t launch_wide_nw(int family, const mx_args& a, hipStream_t s) {
  if (a.splits % 8 != 0 || a.Cp % (32 * NW) != 0 || !a.Rh || !a.Rl) return hipErrorInvalidValue;
  dim3 grid(a.splits * (a.Cp / (32 * NW)));
  if (family == SX_LOGIT)
    hipLaunchKernelGGL((glm_mfma32_fwd<KQ, SX_LOGIT, NW>), grid, dim3(64 * NW), 0, s, a);
  else
    hipLaunchKernelGGL((glm_mfma32_fwd<KQ, SX_NORMAL, NW>), grid, dim3(64 * NW), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((glm_mfma32_adj<8, NW>), grid, dim3(64 * NW), 0, s, a, 0);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((glm_mfma32_adj<KQ - 8, NW>), grid, dim3(64 * NW), 0, s, a, 8);
  return hipGetLastError();
}
template <int KQ>
hipError_t launch_wide(int family, const mx_args& a, hipStream_t s) {
  if (a.Cp == 64) return launch_wide_nw<KQ, 2>(family, a, s);
  return launch_wide_nw<KQ, 4>(family, a, s);
}

}  // namespace

hipError_t launch_split_x(const double* Xrm, uint16_t* X1h, uint16_t* X1l, uint16_t* X2h,
                          uint16_t* X2l, int64_t N, int64_t K, int64_t Np, int64_t Kp,
                          hipStream_t s) {
  const int64_t total = Np * Kp;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(split_x, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s,
                     Xrm, reinterpret_cast<__bf16*>(X1h), reinterpret_cast<__bf16*>(X1l),
                     reinterpret_cast<__bf16*>(X2h), reinterpret_cast<__bf16*>(X2l), N, K, Kp,
                     total);
  return hipGetLastError();
}

hipError_t launch_mfma_pass(int family, const double* theta, int64_t D, int C, const mx_args& a,
                            const pass_args& pa, hipStream_t s) {
  hipLaunchKernelGGL(prep_chains, dim3(a.Cp), dim3(256), 0, s, theta, D, C, pa.K, a.Kp,
                     reinterpret_cast<__bf16*>(const_cast<uint16_t*>(a.Bh)),
                     reinterpret_cast<__bf16*>(const_cast<uint16_t*>(a.Bl)),
                     const_cast<float*>(a.alpha));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // chains come padded to 64 or a multiple of 128, K to a multiple of 32
  // (of 64 above 256)
  if (!(a.Cp == 64 || a.Cp % 128 == 0)) return hipErrorInvalidValue;
  switch (a.Kp / 32) {
    case 1: e = launch_kq32<1>(family, a, s); break;
    case 2: e = launch_kq32<2>(family, a, s); break;
    case 3: e = launch_kq32<3>(family, a, s); break;
    case 4: e = launch_kq32<4>(family, a, s); break;
    case 5: e = launch_kq32<5>(family, a, s); break;
    case 6: e = launch_kq32<6>(family, a, s); break;
    case 7: e = launch_kq32<7>(family, a, s); break;
    case 8: e = launch_kq32<8>(family, a, s); break;
    case 10: e = launch_wide<10>(family, a, s); break;
    case 12: e = launch_wide<12>(family, a, s); break;
    case 14: e = launch_wide<14>(family, a, s); break;
    case 16: e = launch_wide<16>(family, a, s); break;
    default: return hipErrorInvalidValue;
  }
  if (e != hipSuccess) return e;
  const size_t lds = static_cast<size_t>(pa.E) * sizeof(double);
  if (family == SX_LOGIT)
    hipLaunchKernelGGL((mx_reduce<SX_LOGIT>), dim3(C), dim3(1024), lds, s, a, pa);
  else
    hipLaunchKernelGGL((mx_reduce<SX_NORMAL>), dim3(C), dim3(1024), lds, s, a, pa);
  return hipGetLastError();
}

}  // namespace sx
