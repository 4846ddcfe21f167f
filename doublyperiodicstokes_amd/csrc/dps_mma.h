// dps_mma.h -- the 16x16x4 matrix-core step in both grid precisions.
#pragma once
#include <hip/hip_runtime.h>

namespace dps {

// v_mfma 16x16x4: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16];
// output register v of lane l holds D[row(l, v)][l % 16], row 4 (l / 16) + v
// for FP32 and l / 16 + 4 v for FP64 (measured on gfx950).
template <class T> struct Mma;
template <> struct Mma<float> {
  typedef float acc __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc step(float a, float b, acc c)
  {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <> struct Mma<double> {
  typedef double acc __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc step(double a, double b, acc c)
  {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

}  // namespace dps
