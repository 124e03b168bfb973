// Device helpers shared by the gfx950 kernels: complex arithmetic on
// double2 / float2 (re, im), conversions between storage and arithmetic
// types, and the XCD-aware workgroup order.
#pragma once

#include <hip/hip_runtime.h>

namespace td1c {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ void cfma(double2& acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}

// complex64 counterparts, for the mixed-precision mean field
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, double s) {
  const float f = (float)s;
  return make_float2(a.x * f, a.y * f);
}
__device__ __forceinline__ void cfma(float2& acc, float2 a, float2 b) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.x = fmaf(-a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y);
  acc.y = fmaf(a.y, b.x, acc.y);
}

// conversions between the storage type of an array and the arithmetic type
template <class A2> __device__ __forceinline__ A2 cz();
template <> __device__ __forceinline__ double2 cz<double2>() { return make_double2(0.0, 0.0); }
template <> __device__ __forceinline__ float2 cz<float2>() { return make_float2(0.0f, 0.0f); }
template <class A2> __device__ __forceinline__ A2 cto(double2 v);
template <class A2> __device__ __forceinline__ A2 cto(float2 v);
template <> __device__ __forceinline__ double2 cto<double2>(double2 v) { return v; }
template <> __device__ __forceinline__ double2 cto<double2>(float2 v) { return make_double2(v.x, v.y); }
template <> __device__ __forceinline__ float2 cto<float2>(double2 v) { return make_float2((float)v.x, (float)v.y); }
template <> __device__ __forceinline__ float2 cto<float2>(float2 v) { return v; }
// real scalar type of a complex type
template <class A2> struct CReal;
template <> struct CReal<double2> { using T = double; };
template <> struct CReal<float2> { using T = float; };

// Workgroups are dealt to the 8 XCDs round-robin by linear id, and each XCD
// has its own L2.  This maps the hardware linear id to a logical tile id so
// that every XCD gets one contiguous range of logical tiles: tiles that
// re-read the same input then share an L2.
__device__ __forceinline__ int xcd_remap(int lin, int n) {
  const int x = lin & 7, j = lin >> 3, q = n >> 3, r = n & 7;
  return x * q + min(x, r) + j;
}
// The inverse of xcd_remap: the linear id whose logical tile is `job`.
__device__ __forceinline__ int xcd_unmap(int job, int n) {
  const int q = n >> 3, r = n & 7, head = r * (q + 1);
  if (job < head) return (job % (q + 1)) * 8 + job / (q + 1);
  const int d = job - head;
  return (d % q) * 8 + r + d / q;
}

}  // namespace td1c
