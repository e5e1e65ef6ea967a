// FP64 -> three bf16 pieces hi = bf16(x), mid = bf16(x - hi), lo = bf16(x - hi - mid)
// (split3 returns them in bits 0-15, 16-31 and 32-47 of one word), stored by the
// split-BF16x3 engine as three planes; FP64 -> FP32 hi/lo pairs (split2).
#pragma once
#include <hip/hip_runtime.h>

namespace pamd {

// round-to-nearest-even FP32 -> bf16 bits
__device__ __forceinline__ unsigned int bf16_bits(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return u >> 16;
}

__device__ __forceinline__ double bf16_value(unsigned int b) {
  return (double)__uint_as_float(b << 16);
}

__device__ __forceinline__ unsigned long long split3(double x) {
  const unsigned int h = bf16_bits((float)x);
  const double r1 = x - bf16_value(h);
  const unsigned int m = bf16_bits((float)r1);
  const double r2 = r1 - bf16_value(m);
  const unsigned int l = bf16_bits((float)r2);
  return (unsigned long long)h | ((unsigned long long)m << 16) | ((unsigned long long)l << 32);
}

// FP64 -> FP32 hi/lo pair in one 8-byte element: bits 0-31 hi = fp32(x),
// 32-63 lo = fp32(x - hi) (the comparison form of the FP32 MFMA engine).
struct F2 {
  unsigned long long u;
};

__device__ __forceinline__ F2 split2(double x) {
  const float h = (float)x;
  const float l = (float)(x - (double)h);
  F2 r;
  r.u = (unsigned long long)__float_as_uint(h) | ((unsigned long long)__float_as_uint(l) << 32);
  return r;
}

// a planar split-BF16x3 array: element t of plane q at p[t + q * ps]
struct Planar3 {
  unsigned short *p;
  long long ps;
};

__device__ __forceinline__ void put3(unsigned short *p, long long ps, size_t t, double x) {
  const unsigned long long w = split3(x);
  p[t] = (unsigned short)w;
  p[t + ps] = (unsigned short)(w >> 16);
  p[t + 2 * ps] = (unsigned short)(w >> 32);
}

// store policy of the layout kernels: FP64 copy, planar split-BF16x3 or packed FP32 pair
__device__ __forceinline__ void put(double *p, size_t t, double x) { p[t] = x; }
__device__ __forceinline__ void put(Planar3 p, size_t t, double x) { put3(p.p, p.ps, t, x); }
__device__ __forceinline__ void put(F2 *p, size_t t, double x) { p[t] = split2(x); }

}  // namespace pamd
