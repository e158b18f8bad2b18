// gfx950 matrix-core helpers.
//
// v_mfma_f64_16x16x4f64 (wave64): lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; accumulator register v of lane l holds D[l/16 + 4v][l%16]
// (rows interleaved by 4).  Layout measured on MI355X with a unique-value
// probe (D[i][j] = i + 100 j).
#pragma once
#include <hip/hip_runtime.h>

namespace hmat {
namespace mfma {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ inline v4d f64_16x16x4(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline int f64_row(int lane, int v) { return (lane >> 4) + 4 * v; }
__device__ inline int f64_col(int lane, int) { return lane & 15; }

// v_mfma_f32_16x16x4f32: same input layout as the FP64 form; the FP32
// accumulator follows the 32-bit 16x16 layout, register v of lane l holding
// D[4(l/16)+v][l%16].
__device__ inline v4f f32_16x16x4(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ inline int f32_row(int lane, int v) { return 4 * (lane >> 4) + v; }

// v_mfma_f32_16x16x32_bf16: lane l supplies A[l%16][8(l/16)+j] and
// B[8(l/16)+j][l%16] for j = 0..7; accumulator register v of lane l holds
// D[4(l/16)+v][l%16] (rows blocked by 4; same probe).
__device__ inline v4f bf16_16x16x32(v8bf a, v8bf b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ inline int bf16_row(int lane, int v) { return 4 * (lane >> 4) + v; }
__device__ inline int bf16_col(int lane, int) { return lane & 15; }

}  // namespace mfma
}  // namespace hmat
