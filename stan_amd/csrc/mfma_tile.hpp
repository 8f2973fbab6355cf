// Shared pieces of the BF16 matrix-core passes (gfx950): operand vector
// types, the staged-tile LDS layout with its XOR permutation, and the
// transposing LDS read that feeds a contraction over rows.
#ifndef STANX_MFMA_TILE_HPP
#define STANX_MFMA_TILE_HPP

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include "tile_layout.hpp"

namespace sx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void* lds_ptr;  // LDS pointer of global_load_lds
typedef __attribute__((address_space(3))) bf16x4* lds_bf16x4;

__device__ __forceinline__ __bf16 to_bf16(float x) { return static_cast<__bf16>(x); }

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
