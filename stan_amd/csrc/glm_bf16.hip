// Batched GLM log density + gradient for many chains on MFMA (gfx950),
// X held as a BF16 hi+lo split with FP32 accumulation.
//
// For C chains with coefficients B (K x C) the forward pass is the dense
// contraction H = X B and the adjoint is G = X^T R with R the residual
// matrix: two GEMMs per evaluation, which is where MFMA pays (the single-
// chain pass is a matrix-vector product and stays on the HBM roofline).
// All products run on v_mfma_f32_32x32x16_bf16.
//
// Precision.  X = Xh + Xl (two BF16 terms, about 16 significant bits).
// The chains' coefficients carry three BF16 terms, B = Bh + Bl + Bm (the
// FP32 value exactly), because an error in B is the same for every row:
// it does not average out over the N rows the way the per-row errors of X
// and R do, and with two terms it sets the gradient error (about 1e-4
// relative near the posterior mode at N = 1e6).  The forward therefore
// forms Xh Bh + Xh Bl + Xl Bh + Xh Bm, the adjoint Rh Xh + Rh Xl + Rl Xh
// (R = Rh + Rl); the dropped terms (Xl Bl, Xl Bm, Rl Xl) are below the
// representation error of the two-term splits.
//
// Layout.  X is stored once, k-packed (row n: Kp contiguous BF16 values),
// hi and lo in separate arrays.  The forward operand takes 8 consecutive k
// of one row per lane (one 16-byte LDS read); the adjoint contracts over
// rows and takes 4 consecutive rows of one column per lane with gfx950's
// transposing LDS read (ds_read_b64_tr_b16), from the same staged tile.
// In LDS, 16-byte chunk c of tile row r sits at c ^ 2 (r mod 8) (a smaller
// XOR block for short rows), which spreads both access patterns over the
// banks.
//
// Workgroup = NW waves x 32 chains (NW = 4 for 128-chain tiles, 2 for
// batches of at most 64 chains); each 32-row tile of X is staged once per
// workgroup, straight from global memory into LDS (global_load_lds), one
// tile ahead in the other of two buffers.  1-D grid of row splits x chain
// blocks, mapped so that the blocks sharing a row range land on the same
// XCD and share X in its L2.  Per-block FP32 partials are summed across
// blocks in FP64 by a second launch that also adds priors and Jacobian
// (shared finalize).
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cmath>
#include <cstdint>

#include "glm_finalize.hpp"
#include "mfma_tile.hpp"
#include "sx_internal.hpp"

namespace sx {

namespace {


__device__ __forceinline__ void split(float x, __bf16* hi, __bf16* lo) {
  const __bf16 h = to_bf16(x);
  *hi = h;
  *lo = to_bf16(x - static_cast<float>(h));
}

// One (row, chain) of the link, FP32.  Logistic: the chain's coefficients
// and intercept arrive scaled by log2(e) (prep_chains), so e2 = log2(e) eta
// and, with the row's sign s = 2y - 1 and q = s e2,
//   L = log2(1 + 2^-|q|),  log density = -ln2 (L + (|q| - q) / 2),
//   r = s inv_logit(-s eta) = s 2^-(L + (q + |q|) / 2)
// on three hardware transcendentals (v_exp_f32, v_log_f32, v_exp_f32; the
// exp2 arguments are <= 0 and the log2 argument lies in [1, 2], where the
// instructions need none of the library calls' range handling).  |s| L is
// added to l2 and |q| - q to m2 (both halves of max(+-q, 0) are exact), so a
// row past N (s = 0) adds nothing, and a non-finite q makes m2 non-finite
// (|q| - q is NaN or +inf exactly when q is): the caller flags a non-finite
// log density.  Normal: r = y - eta (0 past N), l2 += r^2, and the
// non-finite flag nf is a sum of 0 * eta.
template <int FAM>
__device__ __forceinline__ float link_row(float h, float alpha, float yv, bool live, float& l2,
                                          float& m2, float& nf) {
  if (FAM == SX_LOGIT) {
    const float q = yv * (h + alpha);
    const float L = __builtin_amdgcn_logf(1.0f + __builtin_amdgcn_exp2f(-fabsf(q)));
    l2 = fmaf(fabsf(yv), L, l2);
    m2 += fabsf(q) - q;
    return yv * __builtin_amdgcn_exp2f(-fmaf(0.5f, q + fabsf(q), L));
  } else {
    const float eta = h + alpha;
    const float rr = live ? yv - eta : 0.f;
    nf = fmaf(eta, 0.f, nf);
    l2 = fmaf(rr, rr, l2);
    return rr;
  }
}

// A tile's row sums as the log-density contribution the reductions expect:
// the logistic log density (natural units); the normal sum of squares.
template <int FAM>
__device__ __forceinline__ double tile_lp(float l2, float m2) {
  if (FAM == SX_LOGIT)
    return -0.69314718055994530942 * (static_cast<double>(l2) + 0.5 * static_cast<double>(m2));
  return static_cast<double>(l2);
}

// The pass's status word: 1 when a row's linear predictor (or the log
// density it gives) is not finite.
template <int FAM>
__device__ __forceinline__ double bad_flag(double lp, float nf) {
  if (FAM == SX_LOGIT) return isfinite(lp) ? 0.0 : 1.0;
  return nf == nf ? 0.0 : 1.0;
}

// Row-major FP64 X -> k-packed BF16 hi and lo (Np x Kp, zero padded).
__global__ void split_x(const double* __restrict__ Xrm, __bf16* Xh, __bf16* Xl, int64_t N,
                        int64_t K, int64_t Kp, int64_t total) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int64_t n = idx / Kp;
  const int64_t k = idx - n * Kp;
  const float x = (n < N && k < K) ? static_cast<float>(Xrm[n * K + k]) : 0.0f;
  __bf16 h, l;
  split(x, &h, &l);
  Xh[idx] = h;
  Xl[idx] = l;
}

// y -> one float per padded row: the sign +-1 of the logistic model's
// outcome, the normal model's value; 0 past N.
__global__ void prep_y(const void* __restrict__ y, float* yf, int family, int64_t N,
                       int64_t Np) {
  const int64_t n = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= Np) return;
  float v = 0.f;
  if (n < N) {
    if (family == SX_LOGIT)
      v = static_cast<const int8_t*>(y)[n] ? 1.0f : -1.0f;
    else
      v = static_cast<float>(static_cast<const double*>(y)[n]);
  }
  yf[n] = v;
}

// theta (C x D, FP64) -> B as three BF16 terms (Cp x Kp each: the FP32
// value of scale x beta exactly), alpha (Cp, times scale); scale is log2(e)
// for the logistic link (base 2), 1 for the normal one
__global__ void prep_chains(const double* __restrict__ theta, int64_t D, int C, int K, int Kp,
                            double scale, __bf16* Bh, __bf16* Bl, __bf16* Bm, float* alpha) {
  const int c = blockIdx.x;
  for (int k = threadIdx.x; k < Kp; k += blockDim.x) {
    const float b = (c < C && k < K) ? static_cast<float>(scale * theta[c * D + 1 + k]) : 0.0f;
    const __bf16 h = to_bf16(b);
    const float r1 = b - static_cast<float>(h);  // exact: at most 16 significant bits
    const __bf16 l = to_bf16(r1);
    const __bf16 m = to_bf16(r1 - static_cast<float>(l));  // exact: at most 8 bits
    const int64_t o = static_cast<int64_t>(c) * Kp + k;
    Bh[o] = h;
    Bl[o] = l;
    Bm[o] = m;
  }
  if (threadIdx.x == 0) alpha[c] = c < C ? static_cast<float>(scale * theta[c * D]) : 0.0f;
}

// K <= 256 (KQ = Kp / 32 <= 8): workgroup = NW waves x 32 chains.  Per
// 32-row tile a wave forms
//   forward  H (32 rows x 32 chains) = X tile (A, LDS) . B (registers):
//            lane l holds chain l % 32 and rows 8 (j/4) + 4 (l/32) + j%4
//   link     r and the row terms on the VALU (FP32)
//   adjoint  G^T (32 chains x 32 k) += R^T . X: step st contracts the
//            lane's own r[8st .. 8st+7] with the same rows of X
// v_mfma_f32_32x32x16_bf16 operands (lane l, j < 8): A[l % 32][8 (l / 32)
// + j], B[8 (l / 32) + j][l % 32]; result d[j] = D[8 (j / 4) + 4 (l / 32) +
// j % 4][l % 32].
template <int KQ, int FAM, int NW, bool PF>
__device__ __forceinline__ void mx_tile(const mx_args& a, const __bf16* X, const float* Y,
                                        int64_t t, int lane, const bf16x8 (&bh)[2 * KQ],
                                        const bf16x8 (&bl)[2 * KQ], const bf16x8 (&bm)[2 * KQ],
                                        float alpha, f32x16 (&g)[KQ], double& lp, double& sr,
                                        float& nf) {
  constexpr int Kp = 32 * KQ;
  constexpr int S16 = 2 * KQ;
  constexpr int P = xor_block<Kp / 8>();
  constexpr int TB = 64 * Kp;  // bytes of one tile array (hi or lo)
  // the lane's address terms are formed per tile (a few registers) rather
  // than every read address being kept live across the tile loop
  int ln = lane;
  asm volatile("" : "+v"(ln));
  const tile_addr<2 * Kp, P> ad(ln);
  const int h = ln >> 5;
  const char* xb = reinterpret_cast<const char*>(X);
  // forward: with PF the operands of step s + 1 are read while step s's
  // products run (one step of prefetch); without, each step reads its own
  // (fewer registers: the kernel then runs two waves per SIMD, which hide
  // each other's LDS latency).  The products are the same either way.
  f32x16 H;
#pragma unroll
  for (int j = 0; j < 16; ++j) H[j] = 0.f;
  bf16x8 fh = *reinterpret_cast<const bf16x8*>(xb + ad.fwd(0));
  bf16x8 fl = *reinterpret_cast<const bf16x8*>(xb + TB + ad.fwd(0));
#pragma unroll
  for (int s = 0; s < S16; ++s) {
    bf16x8 nh = fh, nl = fl;
    if (!PF && s > 0) {
      fh = *reinterpret_cast<const bf16x8*>(xb + ad.fwd(s));
      fl = *reinterpret_cast<const bf16x8*>(xb + TB + ad.fwd(s));
    }
    if (PF && s + 1 < S16) {
      nh = *reinterpret_cast<const bf16x8*>(xb + ad.fwd(s + 1));
      nl = *reinterpret_cast<const bf16x8*>(xb + TB + ad.fwd(s + 1));
    }
    H = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fh, bh[s], H, 0, 0, 0);
    H = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fh, bl[s], H, 0, 0, 0);
    H = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fl, bh[s], H, 0, 0, 0);
    H = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fh, bm[s], H, 0, 0, 0);
    if (PF) {
      fh = nh;
      fl = nl;
    }
  }
  // adjoint operands of (st, kb) = (0, 0), read while the link runs (PF)
  bf16x8 ah, al;
  if (PF) {
    ah = tr_operand(xb, ad.tr(0, 0, 0), ad.tr(0, 1, 0));
    al = tr_operand(xb + TB, ad.tr(0, 0, 0), ad.tr(0, 1, 0));
  }
  const int64_t n0 = t * 32;
  float l2 = 0.f, m2 = 0.f, srt = 0.f;
#pragma unroll
  for (int st = 0; st < 2; ++st) {
    bf16x8 rh, rl;
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
      const int j = 8 * st + jj;
      const int row = 8 * (j >> 2) + 4 * h + (j & 3);
      const float yv = Y[row];
      const float rr = link_row<FAM>(H[j], alpha, yv, n0 + row < a.N, l2, m2, nf);
      srt += rr;
      __bf16 hh, ll;
      split(rr, &hh, &ll);
      rh[jj] = hh;
      rl[jj] = ll;
    }
    // (PF) one (st, kb) of prefetch: the operands of the next column block
    // (or of the next half's first) are read while this block's products run
#pragma unroll
    for (int kb = 0; kb < KQ; ++kb) {
      if (!PF) {
        ah = tr_operand(xb, ad.tr(st, 0, kb), ad.tr(st, 1, kb));
        al = tr_operand(xb + TB, ad.tr(st, 0, kb), ad.tr(st, 1, kb));
      }
      bf16x8 nh = ah, nl = al;
      const int nst = kb + 1 < KQ ? st : st + 1, nkb = kb + 1 < KQ ? kb + 1 : 0;
      if (PF && nst < 2) {
        nh = tr_operand(xb, ad.tr(nst, 0, nkb), ad.tr(nst, 1, nkb));
        nl = tr_operand(xb + TB, ad.tr(nst, 0, nkb), ad.tr(nst, 1, nkb));
      }
      g[kb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh, ah, g[kb], 0, 0, 0);
      g[kb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh, al, g[kb], 0, 0, 0);
      g[kb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rl, ah, g[kb], 0, 0, 0);
      if (PF) {
        ah = nh;
        al = nl;
      }
    }
  }
  lp += tile_lp<FAM>(l2, m2);
  sr += static_cast<double>(srt);
}

template <int KQ, int FAM, int NW, bool PF>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(PF ? 1 : 2))) void glm_mfma32(mx_args a) {
  static_assert(KQ >= 1 && KQ <= 8, "glm_mfma32 handles 32 <= Kp <= 256");
  static_assert(NW == 2 || NW == 4, "2 or 4 waves per workgroup");
  constexpr int Kp = 32 * KQ;
  constexpr int S16 = 2 * KQ;          // forward k-steps of 16
  constexpr int CPR = Kp / 8;          // 16-byte chunks per row
  constexpr int P = xor_block<CPR>();
  constexpr int TE = 32 * Kp;          // elements of one tile array
  constexpr int BPA = Kp / 16;         // 1 KB blocks per tile array
  constexpr int NBW = 2 * BPA / NW;    // staging blocks per wave and tile (hi + lo)
  // two tile buffers, hi then lo; separate objects, so a staging write into
  // one is not taken to alias the reads of the other
  __shared__ __attribute__((aligned(16))) __bf16 xa[2 * TE];
  __shared__ __attribute__((aligned(16))) __bf16 xb[2 * TE];
  __shared__ __attribute__((aligned(16))) float ya[32];
  __shared__ __attribute__((aligned(16))) float yb[32];
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
  bf16x8 bh[S16], bl[S16], bm[S16];
#pragma unroll
  for (int s = 0; s < S16; ++s) {
    const int64_t off = static_cast<int64_t>(chain) * Kp + 16 * s + 8 * h;
    bh[s] = *reinterpret_cast<const bf16x8*>(a.Bh + off);
    bl[s] = *reinterpret_cast<const bf16x8*>(a.Bl + off);
    bm[s] = *reinterpret_cast<const bf16x8*>(a.Bm + off);
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
  const __bf16* Xh = reinterpret_cast<const __bf16*>(a.Xh);
  const __bf16* Xl = reinterpret_cast<const __bf16*>(a.Xl);
  // Staging of tile T into buffer XB / YB: wave w moves 1 KB blocks w,
  // w + NW, ... of the tile (hi blocks, then lo), lane l 16 bytes at block
  // offset 16 l, reading chunk swz(r, p % CPR) of row r = p / CPR
  // for LDS chunk p; 8 lanes of wave 0 move the tile's 32 y values.
#define SX_STAGE(T, XB, YB)                                                      \
  {                                                                              \
    const int64_t base_ = (T) * 32 * Kp;                                         \
    int ln_ = lane; /* recomputed per stage, not kept live */                    \
    asm volatile("" : "+v"(ln_));                                                \
    _Pragma("unroll") for (int i = 0; i < NBW; ++i) {                            \
      const int b_ = w + NW * i;                                                 \
      const int bb_ = b_ % BPA;                                                  \
      const int p_ = bb_ * 64 + ln_;                                             \
      const int r_ = p_ / CPR;                                                   \
      const int src_ = r_ * CPR + swz<P>(r_, p_ % CPR);                           \
      __builtin_amdgcn_global_load_lds((b_ < BPA ? Xh : Xl) + base_ + 8 * src_,  \
          (lds_ptr)((XB) + (b_ / BPA) * TE + bb_ * 512), 16, 0, 0);              \
    }                                                                            \
    if (w == 0 && lane < 8)                                                      \
      __builtin_amdgcn_global_load_lds(a.yf + (T) * 32 + 4 * lane, (lds_ptr)(YB), \
                                       16, 0, 0);                                \
  }
  if (t0 < t1) SX_STAGE(t0, xa, ya);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int64_t t = t0; t < t1; t += 2) {
    // xb held tile t - 1, read before the last barrier
    if (t + 1 < t1) SX_STAGE(t + 1, xb, yb);
    mx_tile<KQ, FAM, NW, PF>(a, xa, ya, t, lane, bh, bl, bm, alpha, g, lp, sr, nf);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (t + 1 >= t1) break;
    if (t + 2 < t1) SX_STAGE(t + 2, xa, ya);
    mx_tile<KQ, FAM, NW, PF>(a, xb, yb, t + 1, lane, bh, bl, bm, alpha, g, lp, sr, nf);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
#undef SX_STAGE
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
  const double bad = bad_flag<FAM>(lp, nf);
  if (h == 0) {
    double* sp = a.spart + (pbase + chain) * 3;
    sp[0] = lp;
    sp[1] = sr;
    sp[2] = bad;
  }
}

// 128 < K <= 256, KQ even: the coefficients of a 32-chain block over the
// full K (three terms) and its gradient accumulators do not fit one wave's
// registers beside the operands, so each chain block is split over a pair
// of waves by k: wave (cb, kh) holds the coefficients and accumulates the
// gradient of columns [kh Kp/2, (kh+1) Kp/2).  Per tile both waves form
// their half of the forward product, exchange the 32 x 32 partials through
// LDS (each then holds the full H; own + other is the same sum in both),
// run the link on all 32 rows (the lp / residual sums are kept by kh = 0
// only), and contract R with their half of the columns.  Workgroup = NCB
// chain blocks x 2 halves (NCB = 4: 8 waves, two per SIMD; NCB = 2 for
// batches of at most 64 chains).
template <int KQ, int FAM, int NCB>
__global__ __launch_bounds__(128 * NCB) void glm_mfma32_pair(mx_args a) {
  static_assert(KQ % 2 == 0 && KQ >= 2 && KQ <= 8, "pairs for Kp <= 256, KQ even");
  static_assert(NCB == 2 || NCB == 4, "2 or 4 chain blocks per workgroup");
  constexpr int NW = 2 * NCB;          // waves per workgroup
  constexpr int KH = KQ / 2;           // 32-column blocks per half
  constexpr int Kp = 32 * KQ;
  constexpr int CPR = Kp / 8;
  constexpr int P = xor_block<CPR>();
  constexpr int TE = 32 * Kp;
  constexpr int BPA = Kp / 16;
  constexpr int NBW = 2 * BPA / NW;
  static_assert((2 * BPA) % NW == 0, "staging blocks per wave");
  __shared__ __attribute__((aligned(16))) __bf16 xa[2 * TE];
  __shared__ __attribute__((aligned(16))) __bf16 xb[2 * TE];
  __shared__ __attribute__((aligned(16))) float ya[32];
  __shared__ __attribute__((aligned(16))) float yb[32];
  // forward partials: [chain block][half][4 x 4 elements][lane]
  __shared__ __attribute__((aligned(16))) float4 hx[NCB][2][4][64];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int cbw = w >> 1, kh = w & 1;
  const int li = lane & 31, h = lane >> 5;
  const int nCB = a.Cp / (32 * NCB);
  const int bid = blockIdx.x;
  const int xcd = bid & 7, widx = bid >> 3;
  const int spl = (widx / nCB) * 8 + xcd;
  const int cblk = widx % nCB;
  const int cbase = cblk * 32 * NCB + 32 * cbw;  // first chain of this pair
  const int chain = cbase + li;
  const int s0 = kh * KQ;                   // first k-step of 16 of this half
  bf16x8 bh[KQ], bl[KQ], bm[KQ];
#pragma unroll
  for (int s = 0; s < KQ; ++s) {
    const int64_t off = static_cast<int64_t>(chain) * Kp + 16 * (s0 + s) + 8 * h;
    bh[s] = *reinterpret_cast<const bf16x8*>(a.Bh + off);
    bl[s] = *reinterpret_cast<const bf16x8*>(a.Bl + off);
    bm[s] = *reinterpret_cast<const bf16x8*>(a.Bm + off);
  }
  const float alpha = a.alpha[chain];
  f32x16 g[KH];
#pragma unroll
  for (int i = 0; i < KH; ++i)
#pragma unroll
    for (int j = 0; j < 16; ++j) g[i][j] = 0.f;
  double lp = 0.0, sr = 0.0;
  float nf = 0.f;
  const int64_t t0 = static_cast<int64_t>(spl) * a.tiles_per_split;
  const int64_t t1 = t0 + a.tiles_per_split < a.n_tiles ? t0 + a.tiles_per_split : a.n_tiles;
  const __bf16* Xh = reinterpret_cast<const __bf16*>(a.Xh);
  const __bf16* Xl = reinterpret_cast<const __bf16*>(a.Xl);
#define SX_STAGE(T, XB, YB)                                                      \
  {                                                                              \
    const int64_t base_ = (T) * 32 * Kp;                                         \
    int ln_ = lane; /* recomputed per stage, not kept live */                    \
    asm volatile("" : "+v"(ln_));                                                \
    _Pragma("unroll") for (int i = 0; i < NBW; ++i) {                            \
      const int b_ = w + NW * i;                                                 \
      const int bb_ = b_ % BPA;                                                  \
      const int p_ = bb_ * 64 + ln_;                                             \
      const int r_ = p_ / CPR;                                                   \
      const int src_ = r_ * CPR + swz<P>(r_, p_ % CPR);                           \
      __builtin_amdgcn_global_load_lds((b_ < BPA ? Xh : Xl) + base_ + 8 * src_,  \
          (lds_ptr)((XB) + (b_ / BPA) * TE + bb_ * 512), 16, 0, 0);              \
    }                                                                            \
    if (w == 0 && lane < 8)                                                      \
      __builtin_amdgcn_global_load_lds(a.yf + (T) * 32 + 4 * lane, (lds_ptr)(YB), \
                                       16, 0, 0);                                \
  }
  auto tile = [&](const __bf16* X, const float* Y, int64_t t) {
    int ln = lane;
    asm volatile("" : "+v"(ln));
    const tile_addr<2 * Kp, P> ad(ln);
    const char* xb = reinterpret_cast<const char*>(X);
    constexpr int TB = 64 * Kp;
    f32x16 H;
#pragma unroll
    for (int j = 0; j < 16; ++j) H[j] = 0.f;
    bf16x8 fh = *reinterpret_cast<const bf16x8*>(xb + ad.fwd(s0));
    bf16x8 fl = *reinterpret_cast<const bf16x8*>(xb + TB + ad.fwd(s0));
#pragma unroll
    for (int s = 0; s < KQ; ++s) {
      bf16x8 nh = fh, nl = fl;
      if (s + 1 < KQ) {
        nh = *reinterpret_cast<const bf16x8*>(xb + ad.fwd(s0 + s + 1));
        nl = *reinterpret_cast<const bf16x8*>(xb + TB + ad.fwd(s0 + s + 1));
      }
      H = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fh, bh[s], H, 0, 0, 0);
      H = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fh, bl[s], H, 0, 0, 0);
      H = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fl, bh[s], H, 0, 0, 0);
      H = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fh, bm[s], H, 0, 0, 0);
      fh = nh;
      fl = nl;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
      hx[cbw][kh][q][lane] = make_float4(H[4 * q], H[4 * q + 1], H[4 * q + 2], H[4 * q + 3]);
    // the adjoint's first operands, read across the exchange
    bf16x8 ah = tr_operand(xb, ad.tr(0, 0, kh * KH), ad.tr(0, 1, kh * KH));
    bf16x8 al = tr_operand(xb + TB, ad.tr(0, 0, kh * KH), ad.tr(0, 1, kh * KH));
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 o = hx[cbw][kh ^ 1][q][lane];
      H[4 * q] += o.x;
      H[4 * q + 1] += o.y;
      H[4 * q + 2] += o.z;
      H[4 * q + 3] += o.w;
    }
    const int64_t n0 = t * 32;
    float l2 = 0.f, m2 = 0.f, srt = 0.f;
#pragma unroll
    for (int st = 0; st < 2; ++st) {
      bf16x8 rh, rl;
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const int j = 8 * st + jj;
        const int row = 8 * (j >> 2) + 4 * h + (j & 3);
        const float yv = Y[row];
        const float rr = link_row<FAM>(H[j], alpha, yv, n0 + row < a.N, l2, m2, nf);
        srt += rr;
        __bf16 hh, ll;
        split(rr, &hh, &ll);
        rh[jj] = hh;
        rl[jj] = ll;
      }
#pragma unroll
      for (int kb = 0; kb < KH; ++kb) {
        bf16x8 nh = ah, nl = al;
        const int nst = kb + 1 < KH ? st : st + 1, nkb = kb + 1 < KH ? kb + 1 : 0;
        if (nst < 2) {
          nh = tr_operand(xb, ad.tr(nst, 0, kh * KH + nkb), ad.tr(nst, 1, kh * KH + nkb));
          nl = tr_operand(xb + TB, ad.tr(nst, 0, kh * KH + nkb), ad.tr(nst, 1, kh * KH + nkb));
        }
        g[kb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh, ah, g[kb], 0, 0, 0);
        g[kb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh, al, g[kb], 0, 0, 0);
        g[kb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rl, ah, g[kb], 0, 0, 0);
        ah = nh;
        al = nl;
      }
    }
    if (kh == 0) {
      lp += tile_lp<FAM>(l2, m2);
      sr += static_cast<double>(srt);
    }
  };
  if (t0 < t1) SX_STAGE(t0, xa, ya);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int64_t t = t0; t < t1; t += 2) {
    if (t + 1 < t1) SX_STAGE(t + 1, xb, yb);
    tile(xa, ya, t);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (t + 1 >= t1) break;
    if (t + 2 < t1) SX_STAGE(t + 2, xa, ya);
    tile(xb, yb, t + 1);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
#undef SX_STAGE
  const int64_t pbase = static_cast<int64_t>(spl) * a.Cp;
#pragma unroll
  for (int kb = 0; kb < KH; ++kb)
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int ci = cbase + 8 * (j >> 2) + 4 * h + (j & 3);
      a.gpart[(pbase + ci) * Kp + 32 * (kh * KH + kb) + li] = g[kb][j];
    }
  if (kh == 0) {
    lp += __shfl_xor(lp, 32, 64);
    sr += __shfl_xor(sr, 32, 64);
    nf += __shfl_xor(nf, 32, 64);
    const double bad = bad_flag<FAM>(lp, nf);
    if (h == 0) {
      double* sp = a.spart + (pbase + chain) * 3;
      sp[0] = lp;
      sp[1] = sr;
      sp[2] = bad;
    }
  }
}

// 256 < K <= 512 (KQ = Kp / 32 in {10, 12, 14, 16}): B hi/lo of 32 chains
// over the full K fill 256 registers, so the gradient accumulators cannot
// stay beside them; the evaluation is split in two launches, and B keeps
// two BF16 terms here (a third would not fit either).  This one is
// glm_mfma32's forward and link (same workgroup, grid and tile staging):
// it writes the residual r of every row and chain as the BF16 hi/lo MFMA
// operands the adjoint will use, lane by lane (tile t, 32-chain group g,
// half st: the lane's bf16x8 at R[((t * Cp / 32 + g) * 2 + st) * 64 +
// lane]), and the row terms per split.  glm_mfma32_adj then forms
// G^T = R^T X per block of 256 k.
template <int KQ, int FAM, int NW>
__global__ __launch_bounds__(64 * NW) void glm_mfma32_fwd(mx_args a) {
  static_assert(KQ > 8 && KQ <= 16 && KQ % 2 == 0, "glm_mfma32_fwd handles 256 < Kp <= 512, Kp % 64 == 0");
  static_assert(NW == 2 || NW == 4, "2 or 4 waves per workgroup");
  constexpr int Kp = 32 * KQ;
  constexpr int S16 = 2 * KQ;
  constexpr int CPR = Kp / 8;
  constexpr int P = xor_block<CPR>();
  constexpr int TE = 32 * Kp;
  constexpr int BPA = Kp / 16;
  constexpr int NBW = 2 * BPA / NW;
  __shared__ __attribute__((aligned(16))) __bf16 xa[2 * TE];
  __shared__ __attribute__((aligned(16))) __bf16 xb[2 * TE];
  __shared__ __attribute__((aligned(16))) float ya[32];
  __shared__ __attribute__((aligned(16))) float yb[32];
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
  const __bf16* Xh = reinterpret_cast<const __bf16*>(a.Xh);
  const __bf16* Xl = reinterpret_cast<const __bf16*>(a.Xl);
  bf16x8* Rh = reinterpret_cast<bf16x8*>(a.Rh);
  bf16x8* Rl = reinterpret_cast<bf16x8*>(a.Rl);
#define SX_STAGE(T, XB, YB)                                                      \
  {                                                                              \
    const int64_t base_ = (T) * 32 * Kp;                                         \
    int ln_ = lane; /* recomputed per stage, not kept live */                    \
    asm volatile("" : "+v"(ln_));                                                \
    _Pragma("unroll") for (int i = 0; i < NBW; ++i) {                            \
      const int b_ = w + NW * i;                                                 \
      const int bb_ = b_ % BPA;                                                  \
      const int p_ = bb_ * 64 + ln_;                                             \
      const int r_ = p_ / CPR;                                                   \
      const int src_ = r_ * CPR + swz<P>(r_, p_ % CPR);                           \
      __builtin_amdgcn_global_load_lds((b_ < BPA ? Xh : Xl) + base_ + 8 * src_,  \
          (lds_ptr)((XB) + (b_ / BPA) * TE + bb_ * 512), 16, 0, 0);              \
    }                                                                            \
    if (w == 0 && lane < 8)                                                      \
      __builtin_amdgcn_global_load_lds(a.yf + (T) * 32 + 4 * lane, (lds_ptr)(YB), \
                                       16, 0, 0);                                \
  }
  auto tile = [&](const __bf16* X, const float* Y, int64_t t) {
    int ln = lane;
    asm volatile("" : "+v"(ln));
    const tile_addr<2 * Kp, P> ad(ln);
    const char* xb = reinterpret_cast<const char*>(X);
    constexpr int TB = 64 * Kp;
    f32x16 H;
#pragma unroll
    for (int j = 0; j < 16; ++j) H[j] = 0.f;
    bf16x8 fh = *reinterpret_cast<const bf16x8*>(xb + ad.fwd(0));
    bf16x8 fl = *reinterpret_cast<const bf16x8*>(xb + TB + ad.fwd(0));
#pragma unroll
    for (int s = 0; s < S16; ++s) {
      bf16x8 nh = fh, nl = fl;
      if (s + 1 < S16) {
        nh = *reinterpret_cast<const bf16x8*>(xb + ad.fwd(s + 1));
        nl = *reinterpret_cast<const bf16x8*>(xb + TB + ad.fwd(s + 1));
      }
      H = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fh, bh[s], H, 0, 0, 0);
      H = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fh, bl[s], H, 0, 0, 0);
      H = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fl, bh[s], H, 0, 0, 0);
      fh = nh;
      fl = nl;
    }
    const int64_t n0 = t * 32;
    float l2 = 0.f, m2 = 0.f, srt = 0.f;
#pragma unroll
    for (int st = 0; st < 2; ++st) {
      bf16x8 rh, rl;
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const int j = 8 * st + jj;
        const int row = 8 * (j >> 2) + 4 * h + (j & 3);
        const float yv = Y[row];
        const float rr = link_row<FAM>(H[j], alpha, yv, n0 + row < a.N, l2, m2, nf);
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
    lp += tile_lp<FAM>(l2, m2);
    sr += static_cast<double>(srt);
  };
  if (t0 < t1) SX_STAGE(t0, xa, ya);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int64_t t = t0; t < t1; t += 2) {
    if (t + 1 < t1) SX_STAGE(t + 1, xb, yb);
    tile(xa, ya, t);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (t + 1 >= t1) break;
    if (t + 2 < t1) SX_STAGE(t + 2, xa, ya);
    tile(xb, yb, t + 1);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
#undef SX_STAGE
  const int64_t pbase = static_cast<int64_t>(spl) * a.Cp;
  lp += __shfl_xor(lp, 32, 64);
  sr += __shfl_xor(sr, 32, 64);
  nf += __shfl_xor(nf, 32, 64);
  const double bad = bad_flag<FAM>(lp, nf);
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
// 4 h + j%4, j = 8 st .. 8 st + 7, and the same rows of X through the
// transposing LDS read), with R read from the forward launch's output and
// the tile's column block of X staged through LDS one tile ahead.
template <int KQA, int NW>
__global__ __launch_bounds__(64 * NW) void glm_mfma32_adj(mx_args a, int kq0) {
  static_assert(KQA >= 1 && KQA <= 8, "blocks of at most 256 columns");
  static_assert((4 * KQA) % NW == 0, "staging blocks per wave");
  constexpr int KR = 32 * KQA;         // columns of the block
  constexpr int CPRB = KR / 8;         // 16-byte chunks per block row
  constexpr int P = xor_block<CPRB>();
  constexpr int TE = 32 * KR;          // elements of one tile array
  constexpr int BPA = 2 * KQA;         // 1 KB staging blocks per tile array
  constexpr int NBW = 2 * BPA / NW;
  __shared__ __attribute__((aligned(16))) __bf16 xa[2 * TE];
  __shared__ __attribute__((aligned(16))) __bf16 xb[2 * TE];
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
  const __bf16* Xh = reinterpret_cast<const __bf16*>(a.Xh);
  const __bf16* Xl = reinterpret_cast<const __bf16*>(a.Xl);
  const bf16x8* Rh = reinterpret_cast<const bf16x8*>(a.Rh);
  const bf16x8* Rl = reinterpret_cast<const bf16x8*>(a.Rl);
  // staging: LDS chunk p of an array is row p / CPRB of the tile, chunk
  // swz(row, p % CPRB) of the column block
#define SX_STAGE(T, XB)                                                          \
  {                                                                              \
    _Pragma("unroll") for (int i = 0; i < NBW; ++i) {                            \
      const int b_ = w + NW * i;                                                 \
      const int bb_ = b_ % BPA;                                                  \
      const int p_ = bb_ * 64 + lane;                                            \
      const int r_ = p_ / CPRB;                                                  \
      const int64_t src_ = ((T) * 32 + r_) * Kp + k0 + 8 * swz<P>(r_, p_ % CPRB);  \
      __builtin_amdgcn_global_load_lds((b_ < BPA ? Xh : Xl) + src_,              \
          (lds_ptr)((XB) + (b_ / BPA) * TE + bb_ * 512), 16, 0, 0);              \
    }                                                                            \
  }
  auto tile = [&](const __bf16* X, int64_t t) {
    int ln = lane;
    asm volatile("" : "+v"(ln));
    const tile_addr<2 * KR, P> ad(ln);
    const char* xb = reinterpret_cast<const char*>(X);
    constexpr int TB = 64 * KR;
    bf16x8 rh[2], rl[2];
#pragma unroll
    for (int st = 0; st < 2; ++st) {
      const int64_t ri = ((t * nG + g32) * 2 + st) * 64 + lane;
      rh[st] = Rh[ri];
      rl[st] = Rl[ri];
    }
#pragma unroll
    for (int st = 0; st < 2; ++st) {
#pragma unroll
      for (int kb = 0; kb < KQA; ++kb) {
        const bf16x8 xh = tr_operand(xb, ad.tr(st, 0, kb), ad.tr(st, 1, kb));
        const bf16x8 xl = tr_operand(xb + TB, ad.tr(st, 0, kb), ad.tr(st, 1, kb));
        g[kb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh[st], xh, g[kb], 0, 0, 0);
        g[kb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rh[st], xl, g[kb], 0, 0, 0);
        g[kb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(rl[st], xh, g[kb], 0, 0, 0);
      }
    }
  };
  if (t0 < t1) SX_STAGE(t0, xa);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int64_t t = t0; t < t1; t += 2) {
    if (t + 1 < t1) SX_STAGE(t + 1, xb);
    tile(xa, t);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (t + 1 >= t1) break;
    if (t + 2 < t1) SX_STAGE(t + 2, xa);
    tile(xb, t + 1);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
#undef SX_STAGE
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

template <int KQ, int NW, bool PF = true>
hipError_t launch_kq32_nw(int family, const mx_args& a, hipStream_t s) {
  if (a.splits % 8 != 0 || a.Cp % (32 * NW) != 0) return hipErrorInvalidValue;
  dim3 grid(a.splits * (a.Cp / (32 * NW)));
  if (family == SX_LOGIT)
    hipLaunchKernelGGL((glm_mfma32<KQ, SX_LOGIT, NW, PF>), grid, dim3(64 * NW), 0, s, a);
  else
    hipLaunchKernelGGL((glm_mfma32<KQ, SX_NORMAL, NW, PF>), grid, dim3(64 * NW), 0, s, a);
  return hipGetLastError();
}

template <int KQ, int NCB>
hipError_t launch_pair_ncb(int family, const mx_args& a, hipStream_t s) {
  if (a.splits % 8 != 0 || a.Cp % (32 * NCB) != 0) return hipErrorInvalidValue;
  dim3 grid(a.splits * (a.Cp / (32 * NCB)));
  if (family == SX_LOGIT)
    hipLaunchKernelGGL((glm_mfma32_pair<KQ, SX_LOGIT, NCB>), grid, dim3(128 * NCB), 0, s, a);
  else
    hipLaunchKernelGGL((glm_mfma32_pair<KQ, SX_NORMAL, NCB>), grid, dim3(128 * NCB), 0, s, a);
  return hipGetLastError();
}

// Kp <= 256.  Variant 0: one wave per 32-chain block (two waves per
// workgroup for a batch of at most 64 chains (Cp = 64), four otherwise);
// variant 1 (KQ even): a pair of waves per chain block, split by k;
// variant 2 (KQ <= 4): variant 0 without operand prefetch, two waves per
// SIMD (for Cp = 64 it is variant 0: the same products, bit for bit).
template <int KQ>
hipError_t launch_kq32(int family, const mx_args& a, hipStream_t s) {
  if constexpr (KQ % 2 == 0) {
    if (a.variant == 1) {
      if (a.Cp == 64) return launch_pair_ncb<KQ, 2>(family, a, s);
      return launch_pair_ncb<KQ, 4>(family, a, s);
    }
  }
  if constexpr (KQ <= 4) {
    if (a.variant == 2 && a.Cp != 64) return launch_kq32_nw<KQ, 4, false>(family, a, s);
  }
  if (a.variant != 0 && !(a.variant == 2 && KQ <= 4)) return hipErrorInvalidValue;
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

hipError_t launch_split_x(const double* Xrm, uint16_t* Xh, uint16_t* Xl, int64_t N, int64_t K,
                          int64_t Np, int64_t Kp, hipStream_t s) {
  const int64_t total = Np * Kp;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(split_x, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s,
                     Xrm, reinterpret_cast<__bf16*>(Xh), reinterpret_cast<__bf16*>(Xl), N, K, Kp,
                     total);
  return hipGetLastError();
}

hipError_t launch_prep_y(const void* y, float* yf, int family, int64_t N, int64_t Np,
                         hipStream_t s) {
  if (Np == 0) return hipSuccess;
  hipLaunchKernelGGL(prep_y, dim3(static_cast<unsigned>((Np + 255) / 256)), dim3(256), 0, s, y, yf,
                     family, N, Np);
  return hipGetLastError();
}

hipError_t launch_mfma_pass(int family, const double* theta, int64_t D, int C, const mx_args& a,
                            const pass_args& pa, hipStream_t s) {
  hipLaunchKernelGGL(prep_chains, dim3(a.Cp), dim3(256), 0, s, theta, D, C, pa.K, a.Kp,
                     family == SX_LOGIT ? 1.4426950408889634074 : 1.0,
                     reinterpret_cast<__bf16*>(const_cast<uint16_t*>(a.Bh)),
                     reinterpret_cast<__bf16*>(const_cast<uint16_t*>(a.Bl)),
                     reinterpret_cast<__bf16*>(const_cast<uint16_t*>(a.Bm)),
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
