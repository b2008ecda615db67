// MFMA helpers for the pruned-DFT spot kernels (gfx950, wave64).
//
// v_mfma_f32_16x16x32_{f16,bf16}: D(16x16, f32) += A(16x32) * B(32x16).
// Operand layouts used throughout (checked on the device against a plain
// matrix product by the shwfs_mfma_probe entry point and its test):
//   A[i][k]: lane = i + 16*(k/8), vector element k%8
//   B[k][j]: lane = j + 16*(k/8), vector element k%8
//   D[i][j]: lane = j + 16*(i/4), vector element i%4
// Any consistent permutation of k between A and B leaves the product
// unchanged, so only the i / j placement of A, B and D matters.
#pragma once

#include <hip/hip_runtime.h>

namespace shwfs {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma_f16(const f16x8& a, const f16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma_bf16(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// bf16 round-to-nearest-even of a float (finite inputs), and the split
// x = hi + lo with hi = bf16(x), lo = bf16(x - hi).
__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }
__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

}  // namespace shwfs
