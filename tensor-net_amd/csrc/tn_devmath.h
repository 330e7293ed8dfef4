// Device helpers shared by the kernel sources: complex arithmetic on
// double2 and deterministic block reductions.
#pragma once
#include <hip/hip_runtime.h>

namespace tnmps {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 cscale(double2 a, double s) {
  return make_double2(a.x * s, a.y * s);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }
// complex / real as num_complex divides by Complex{re: d, im: 0}
__device__ __forceinline__ double2 cdivr(double2 a, double d) {
  const double den = d * d;
  return make_double2((a.x * d) / den, (a.y * d) / den);
}

// Block-wide sum of one double; deterministic (fixed tree).  `buf` holds at
// least blockDim.x / 64 doubles.  Every thread gets the result.
__device__ __forceinline__ double block_sum(double v, double* buf) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) buf[wid] = v;
  __syncthreads();
  double s = 0.0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int i = 0; i < nw; ++i) s += buf[i];
  __syncthreads();
  return s;
}

__device__ __forceinline__ double block_max(double v, double* buf) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) buf[wid] = v;
  __syncthreads();
  double s = 0.0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int i = 0; i < nw; ++i) s = fmax(s, buf[i]);
  __syncthreads();
  return s;
}

}  // namespace tnmps
