// Counter-based workload generator shared by host code and kernels.
//
// Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): a 4 x 64-bit counter and a 2 x 64-bit key give
// four 64-bit words.  A word w becomes the uniform double (w >> 11) 2^-53 in
// [0, 1).
//
// Streams of one trajectory (key = (seed, 0); counter = (block, c1, layer,
// trajectory id)):
//   c1 = site                 Haar 4x4 unitary of the gate on (site, site+1):
//                             blocks 0..7, 32 uniforms
//   c1 = kMaskStream          measurement mask: qubit q is measured when
//                             uniform q < p (uniform_meas, circuit.rs:580-584)
//   c1 = kDrawStream          Born-sampling draws, one per sampled qubit in
//                             increasing site order
// A Haar unitary is the Q of the QR decomposition, with positive real
// diagonal R, of a 4x4 complex Gaussian matrix Z (Re, Im of each entry
// standard normal, entries row-major; the normals come in Box-Muller pairs
// from consecutive uniforms): the distribution of gate.rs:888-909.
#pragma once

#include <cstdint>

#if defined(__HIP_DEVICE_COMPILE__)
#define TN_HD __host__ __device__
#elif defined(__HIPCC__)
#define TN_HD __host__ __device__
#else
#define TN_HD
#endif

namespace tnmps {

constexpr uint64_t kMaskStream = 1ull << 32;
constexpr uint64_t kDrawStream = 2ull << 32;

struct Philox4x64 {
  uint64_t v[4];
};

TN_HD inline uint64_t mulhilo64(uint64_t a, uint64_t b, uint64_t* hi) {
#if defined(__HIP_DEVICE_COMPILE__)
  *hi = __umul64hi(a, b);
  return a * b;
#else
  const unsigned __int128 p = (unsigned __int128)a * b;
  *hi = (uint64_t)(p >> 64);
  return (uint64_t)p;
#endif
}

// Ten rounds of Philox4x64 on counter c with key (k0, k1).
TN_HD inline Philox4x64 philox4x64_10(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3,
                                      uint64_t k0, uint64_t k1) {
  constexpr uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  constexpr uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
  for (int r = 0; r < 10; ++r) {
    uint64_t hi0, hi1;
    const uint64_t lo0 = mulhilo64(M0, c0, &hi0);
    const uint64_t lo1 = mulhilo64(M1, c2, &hi1);
    const uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += W0;
    k1 += W1;
  }
  Philox4x64 out;
  out.v[0] = c0;
  out.v[1] = c1;
  out.v[2] = c2;
  out.v[3] = c3;
  return out;
}

TN_HD inline double u01(uint64_t w) { return (double)(w >> 11) * 0x1.0p-53; }

}  // namespace tnmps
