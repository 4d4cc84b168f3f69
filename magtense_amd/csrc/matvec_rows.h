// Plain FP32 demag rows, shared by the per-state mat-vec (demag.hip) and the
// fused small-problem kernel (llg.hip), so that both form the same sums in the
// same order.
#pragma once

#include "mt_common.h"

namespace mt {

__device__ inline float4 ldnt4(const float* p) {
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  const f32x4 v = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

// Rows i0 .. i0 + RB - 1 of (K m) for one state (mx, my = mx + n, mz = my + n)
// with a block of 256 threads: every thread sums its columns over the six
// receiver-major planes, then a shuffle tree per wave and the four waves in
// order.  On return (after a barrier) sums[r * 3 + c] holds component c of
// row i0 + r; red is the block's scratch.
template <int RB, bool VEC4>
__device__ inline void matvec_rows_f32(const float* __restrict__ K, int n, const float* __restrict__ mx, int i0,
                                       float (&red)[4][RB * 3], float (&sums)[RB * 3]) {
  const float* my = mx + n;
  const float* mz = my + n;
  const size_t nn = (size_t)n * n;
  float acc[RB][3];
#pragma unroll
  for (int r = 0; r < RB; ++r) acc[r][0] = acc[r][1] = acc[r][2] = 0.0f;
  if (VEC4) {
    for (int j = threadIdx.x * 4; j < n; j += 256 * 4) {
      float4 vx = *reinterpret_cast<const float4*>(mx + j);
      float4 vy = *reinterpret_cast<const float4*>(my + j);
      float4 vz = *reinterpret_cast<const float4*>(mz + j);
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        int i = i0 + r;
        if (i >= n) break;
        // the tensor is streamed once per state: non-temporal loads, so that it
        // does not displace the state vectors from L2
        const float* row = K + (size_t)i * n + j;
        float4 kxx = ldnt4(row);
        float4 kxy = ldnt4(row + nn);
        float4 kxz = ldnt4(row + 2 * nn);
        float4 kyy = ldnt4(row + 3 * nn);
        float4 kyz = ldnt4(row + 4 * nn);
        float4 kzz = ldnt4(row + 5 * nn);
        acc[r][0] += kxx.x * vx.x + kxy.x * vy.x + kxz.x * vz.x + kxx.y * vx.y + kxy.y * vy.y + kxz.y * vz.y +
                     kxx.z * vx.z + kxy.z * vy.z + kxz.z * vz.z + kxx.w * vx.w + kxy.w * vy.w + kxz.w * vz.w;
        acc[r][1] += kxy.x * vx.x + kyy.x * vy.x + kyz.x * vz.x + kxy.y * vx.y + kyy.y * vy.y + kyz.y * vz.y +
                     kxy.z * vx.z + kyy.z * vy.z + kyz.z * vz.z + kxy.w * vx.w + kyy.w * vy.w + kyz.w * vz.w;
        acc[r][2] += kxz.x * vx.x + kyz.x * vy.x + kzz.x * vz.x + kxz.y * vx.y + kyz.y * vy.y + kzz.y * vz.y +
                     kxz.z * vx.z + kyz.z * vy.z + kzz.z * vz.z + kxz.w * vx.w + kyz.w * vy.w + kzz.w * vz.w;
      }
    }
  } else {
    for (int j = threadIdx.x; j < n; j += 256) {
      float vx = mx[j], vy = my[j], vz = mz[j];
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        int i = i0 + r;
        if (i >= n) break;
        const float* row = K + (size_t)i * n + j;
        float kxx = row[0], kxy = row[nn], kxz = row[2 * nn], kyy = row[3 * nn], kyz = row[4 * nn],
              kzz = row[5 * nn];
        acc[r][0] += kxx * vx + kxy * vy + kxz * vz;
        acc[r][1] += kxy * vx + kyy * vy + kyz * vz;
        acc[r][2] += kxz * vx + kyz * vy + kzz * vz;
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float v = acc[r][c];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      if (lane == 0) red[wave][r * 3 + c] = v;
    }
  __syncthreads();
  if (threadIdx.x < RB * 3)
    sums[threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
  __syncthreads();
}

}  // namespace mt
