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

// *p when ok, else 0.  The load issues either way (from `safe`, a valid
// address, when not ok) and the result is masked bitwise with a mask the
// compiler cannot see through, so a run of these issues back to back: as a
// conditional load (which the compiler makes of a plain select), each got a
// branch (and in some kernels a wait for its data) of its own.
__device__ __forceinline__ double2 ld_or_zero(const double2* p, bool ok, const double2* safe) {
  const double2 v = *(ok ? p : safe);
  // one 32-bit mask for all four words (one register live while the load
  // is in flight)
  unsigned m = ok ? ~0u : 0u;
  asm volatile("" : "+v"(m));
  const unsigned long long m2 = ((unsigned long long)m << 32) | m;
  return make_double2(__longlong_as_double(__double_as_longlong(v.x) & (long long)m2),
                      __longlong_as_double(__double_as_longlong(v.y) & (long long)m2));
}

// Cross-lane moves and wave sums on the DPP network (no LDS): CTRL is a
// DPP control word (quad_perm 0x00-0xff, row_half_mirror 0x141, row_mirror
// 0x140).  A double moves as its two 32-bit halves.
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double read_lane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over the 64 lanes of a wave, returned to every lane, in a fixed order
// (xor 1, xor 2, half-row mirror, row mirror: each 16-lane row then holds
// its row sum; the four row sums are added as (r0 + r1) + (r2 + r3)).  Every
// lane of the wave must be active.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_mov<0xB1>(v);
  v += dpp_mov<0x4E>(v);
  v += dpp_mov<0x141>(v);
  v += dpp_mov<0x140>(v);
  return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// Maximum over the 64 lanes of a wave, returned to every lane, by the same
// moves (fmax is exact: any order gives the same value).  Every lane of the
// wave must be active.
__device__ __forceinline__ double wave_max_dpp(double v) {
  v = fmax(v, dpp_mov<0xB1>(v));
  v = fmax(v, dpp_mov<0x4E>(v));
  v = fmax(v, dpp_mov<0x141>(v));
  v = fmax(v, dpp_mov<0x140>(v));
  return fmax(fmax(read_lane(v, 0), read_lane(v, 16)), fmax(read_lane(v, 32), read_lane(v, 48)));
}

// The value of the first lane of the lane's group of 8 (lane l & ~7), on the
// DPP network: a quad broadcast (quad_perm [0,0,0,0]), then row_shr:4 into
// lanes 4-7 of each group only (banks 1 and 3 of each row).  Every lane of
// the wave must be active.
__device__ __forceinline__ int bcast8_first(int x) {
  const int v = __builtin_amdgcn_mov_dpp(x, 0x00, 0xF, 0xF, false);
  return __builtin_amdgcn_update_dpp(v, v, 0x114, 0xF, 0xA, false);
}
__device__ __forceinline__ double bcast8_first(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = bcast8_first((int)b), hi = bcast8_first((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// The same for both parts of a complex value, the two chains interleaved.
__device__ __forceinline__ double2 wave_csum_dpp(double2 v) {
  v.x += dpp_mov<0xB1>(v.x);
  v.y += dpp_mov<0xB1>(v.y);
  v.x += dpp_mov<0x4E>(v.x);
  v.y += dpp_mov<0x4E>(v.y);
  v.x += dpp_mov<0x141>(v.x);
  v.y += dpp_mov<0x141>(v.y);
  v.x += dpp_mov<0x140>(v.x);
  v.y += dpp_mov<0x140>(v.y);
  return make_double2((read_lane(v.x, 0) + read_lane(v.x, 16)) + (read_lane(v.x, 32) + read_lane(v.x, 48)),
                      (read_lane(v.y, 0) + read_lane(v.y, 16)) + (read_lane(v.y, 32) + read_lane(v.y, 48)));
}

// In-place bitonic sort of key[0 .. P2) (P2 a power of two) by decreasing
// key, ties by increasing idx, over the workgroup; every thread takes part
// and the sorted arrays are visible to all of them on return.
__device__ __forceinline__ void bitonic_desc(double* key, int* idx, int P2) {
  for (int k = 2; k <= P2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P2; i += blockDim.x) {
        const int l = i ^ j;
        if (l > i) {
          const bool desc = ((i & k) == 0);
          const double ki = key[i], kl = key[l];
          const int ii = idx[i], il = idx[l];
          // "i before l" in the final order
          const bool i_first = (ki > kl) || (ki == kl && ii < il);
          if (desc ? !i_first : i_first) {
            key[i] = kl; key[l] = ki;
            idx[i] = il; idx[l] = ii;
          }
        }
      }
      __syncthreads();
    }
  }
}

}  // namespace tnmps
