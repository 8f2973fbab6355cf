// Shared pieces of the BF16 matrix-core passes (gfx950): operand vector
// types, the staged-tile LDS layout with its XOR permutation, and the
// transposing LDS read that feeds a contraction over rows.
#ifndef STANX_MFMA_TILE_HPP
#define STANX_MFMA_TILE_HPP

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

namespace sx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void* lds_ptr;  // LDS pointer of global_load_lds
typedef __attribute__((address_space(3))) bf16x4* lds_bf16x4;

__device__ __forceinline__ __bf16 to_bf16(float x) { return static_cast<__bf16>(x); }

// Staged tile layout in LDS: row r of a tile occupies ROWB bytes; its
// 16-byte chunk c (8 consecutive columns) sits at chunk position
// c ^ f(r), f(r) = 2 (r & (P/2 - 1)), P the largest power of two <= 16
// dividing the chunks per row.  The XOR spreads both read patterns over the
// banks: the forward's 16-byte reads (8 consecutive rows, one chunk each)
// and the transposing reads (4 consecutive rows x 2 adjacent chunks per 16
// lanes).  Because f(r) only changes bits 1..3 of the chunk index and
// depends on r mod 8, every read address is a per-lane base, one XOR with
// a per-lane value and an immediate.
template <int ROWB, int P>
struct tile_addr {
  static_assert(P == 4 || P == 8 || P == 16, "XOR block of 4, 8 or 16 chunks");
  static constexpr int GM = P / 2 - 1;
  int fb, fx, tb, tx;
  __device__ __forceinline__ explicit tile_addr(int lane) {
    const int li = lane & 31, h = lane >> 5;
    fb = li * ROWB + 16 * h;
    fx = 32 * (li & GM);
    // transposing reads: within each group of 16 lanes, lane i reads 8
    // bytes of row 4 hh + i / 4 (plus 16 st + 8 q), columns c0 + 16 ch +
    // 4 (i % 4) .. + 3
    const int i = lane & 15, g = lane >> 4, hh = g >> 1, ch = g & 1;
    const int rr = 4 * hh + (i >> 2);
    const int f = 2 * (rr & GM);
    const int cl = 2 * ch + ((i & 3) >> 1);
    tb = rr * ROWB + 16 * (cl ^ (f & 3)) + 8 * (i & 1);
    tx = 16 * (f & ~3);
  }
  // byte offset of the forward operand of k-step s (columns 16 s + 8 h ..)
  __device__ __forceinline__ int fwd(int s) const {
    return fb + ((32 * (s % (P / 2))) ^ fx) + 16 * P * (s / (P / 2));
  }
  // byte offset of the transposing read q of half st, column block kb
  __device__ __forceinline__ int tr(int st, int q, int kb) const {
    return tb + (16 * st + 8 * q) * ROWB + 16 * P * (kb / (P / 4)) +
           ((64 * (kb % (P / 4))) ^ tx);
  }
};

// chunk position of global chunk c of tile row r (and the reverse)
template <int P>
__device__ __forceinline__ int swz(int r, int c) { return c ^ (2 * (r & (P / 2 - 1))); }

template <int CPR>
constexpr int xor_block() { return CPR % 16 == 0 ? 16 : CPR % 8 == 0 ? 8 : 4; }

// The adjoint's X operand for the lane's column (c0 + lane % 32) and rows
// r0 + 8 (j / 4) + 4 (lane / 32) + j % 4 (j < 8): element j of the MFMA B
// operand.  Two transposing reads: lane i of a 16-lane group receives
// column i of the 4 rows the group read (result element e = row e).
__device__ __forceinline__ bf16x8 tr_operand(const char* base, int off0, int off1) {
  const bf16x4 p0 = __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4)(base + off0));
  const bf16x4 p1 = __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4)(base + off1));
  bf16x8 v;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    v[e] = p0[e];
    v[4 + e] = p1[e];
  }
  return v;
}

}  // namespace sx

#endif
