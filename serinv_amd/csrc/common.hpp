// Shared device helpers for the gfx950 (CDNA4) BTA kernels.
//
// Matrix storage everywhere is row-major (C-contiguous) float64, the block
// layout of serinv: element (r, c) of a block with leading dimension ld sits
// at ptr[r * ld + c].
//
// FP64 MFMA v_mfma_f64_16x16x4_f64 (one wave, 64 lanes), lane layout measured
// on gfx950:
//   A operand (16 x 4):  lane l holds A[l % 16][l / 16]
//   B operand (4 x 16):  lane l holds B[l / 16][l % 16]
//   C/D (16 x 16, 4 regs per lane): lane l, reg r holds D[4 * r + l / 16][l % 16]
// FP32 v_mfma_f32_16x16x4_f32 has the same A/B layout but
//   C/D: lane l, reg r holds D[4 * (l / 16) + r][l % 16].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rocbta {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f64x4 mfma_f64(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma_f32(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Row / column of accumulator register r held by lane l (16x16 tile).
__device__ __forceinline__ int acc_row_f64(int lane, int r) { return 4 * r + (lane >> 4); }
__device__ __forceinline__ int acc_row_f32(int lane, int r) { return 4 * (lane >> 4) + r; }
__device__ __forceinline__ int acc_col(int lane) { return lane & 15; }

// FP64 operands into either accumulator: f64x4 runs FP64 MFMA, f32x4 rounds
// the operands to FP32 and runs FP32 MFMA (the mixed-precision mode).
__device__ __forceinline__ f64x4 mfma_acc(double a, double b, f64x4 c) { return mfma_f64(a, b, c); }
__device__ __forceinline__ f32x4 mfma_acc(double a, double b, f32x4 c) {
  return mfma_f32((float)a, (float)b, c);
}
__device__ __forceinline__ int acc_row(const f64x4&, int lane, int r) { return acc_row_f64(lane, r); }
__device__ __forceinline__ int acc_row(const f32x4&, int lane, int r) { return acc_row_f32(lane, r); }

// Record an LAPACK-style failure index (1-based; smallest index wins; 0 = ok).
__device__ __forceinline__ void record_info(int* info, int value) {
  if (!info) return;
  int cur = atomicAdd(info, 0);
  while (cur == 0 || value < cur) {
    int prev = atomicCAS(info, cur, value);
    if (prev == cur) break;
    cur = prev;
  }
}

}  // namespace rocbta
