// HIP kernels of the Schmidt-form MPS update path, written for CDNA4 (gfx950).
//
// Every kernel takes an array of task descriptors and uses blockIdx.y as the
// task index, so one launch serves one update, the independent gates of a
// brickwork layer, or the lock-stepped updates of many trajectories.  Bond
// dimensions are read from device memory: the host never waits for a rank.
//
// Kernels (reference lines each one replaces):
//   k_theta        theta = lambda Gamma lambda Gamma lambda, gate applied
//                  (mps.rs:628-651, 732-754, 1343-1352)
//   k_winit        W = I for the Jacobi SVD (the SVD steps, standing in for
//                  zgesvd at mps.rs:299, are in tn_jacobi.hip)
//   k_finish       singular values, sort, rank rule, renormalisation
//                  (process_svd, mps.rs:236-293)
//   k_fmul         factored updates: kept singular vectors times the
//                  isometric factor
//   k_split        new Gamma[k], Gamma[k+1] with lambda divisions and partial
//                  norms (mps.rs:852-876, 1357-1380, local_norm 1147-1202)
//   k_renorm       local_renormalize of both sites (mps.rs:1206-1209)
//   k_measure_part, k_measure_project
//                  Born probabilities, sampling or forced outcome, projection
//                  (mps.rs:1027-1134, 1427-1465, 2229-2242)
//   k_entropy      von Neumann / Renyi entropy of one bond (mps.rs:1251-1282)
//   k_u1           single-site unitary (mps.rs:1298-1314)
//   k_haar4        Haar 4x4 unitaries from the counter-based generator
//                  (tn_rng.h; the distribution of gate.rs:888-909)
#include "tn_device.h"
#include "tn_devmath.h"
#include "tn_kernels.h"
#include "tn_rng.h"

#include <algorithm>
#include <cfloat>

namespace tnmps {

// ---------------------------------------------------------------- theta
// Shape of a task's SVD problem.  The factored modes apply when requested
// (no gate) and K < min(m, n); otherwise the full theta is decomposed.  The
// host bounds C by min(K, m, n) for a factored request: C never exceeds it.
struct TaskShape {
  int m, n, K, mode, R, C, orient;
};
__device__ __forceinline__ TaskShape task_shape(const UpdDesc& d) {
  TaskShape t;
  t.m = 2 * d.dims[0];
  t.K = d.dims[1];
  t.n = 2 * d.dims[2];
  t.mode = (d.mode != 0 && !d.has_gate && t.K < min(t.m, t.n)) ? d.mode : 0;
  if (t.mode == 1) {
    t.R = t.n; t.C = t.K; t.orient = 1;
  } else if (t.mode == 2) {
    t.R = t.m; t.C = t.K; t.orient = 0;
  } else {
    t.orient = (t.m >= t.n) ? 0 : 1;
    t.R = max(t.m, t.n);
    t.C = min(t.m, t.n);
  }
  return t;
}

// The contraction theta0 = A . lamM . B is a complex GEMM of
//   A = Gamma[k] as (2 chiL) x chiM, row (v, s)      (its row-major layout)
//   B = Gamma[k+1] as chiM x (2 chiR), column (t, w) (its row-major layout)
// split into real GEMMs on the matrix cores:
//   Re += Ar Br - Ai Bi,  Im += Ar Bi + Ai Br.
// A block computes a TM x TM tile of theta0: rows (v, s) for TM/2 values of v
// and both s, columns (t, w) for TM/2 values of w and both t, so every
// (v, w) point has its four (s, t) values in the block and the gate, lamL and
// lamR are applied in the epilogue.  K is staged through LDS 16 at a time;
// each of the four waves owns a (TM/2) x (TM/2) quarter as 16 x 16 subtiles.
//
// Arithmetic per precision code P (UpdDesc::prec):
//   0  FP64: v_mfma_f64_16x16x4f64, FP64 accumulation
//   1  FP32: v_mfma_f32_16x16x4f32, operands rounded to FP32, FP32 accumulation
//   2  BF16 split: each FP32 operand x = hi + lo with hi, lo in BF16;
//      a b ~ ah bh + ah bl + al bh on v_mfma_f32_16x16x16bf16_1k, FP32
//      accumulation (about 16 bits of each product are kept)
// Lane layouts (16 x 16 output tiles; lane l, accumulator element e):
//   f64 16x16x4:  A[i][k] lane i + 16k, B[k][j] lane j + 16k; D row l/16 + 4e
//   f32 16x16x4:  A, B as for f64;                            D row 4(l/16) + e
//   bf16 16x16x16: A[i][k] lane i + 16(k/4) element k%4, B[k][j] lane
//                 j + 16(k/4) element k%4;                    D row 4(l/16) + e
// and D column l % 16 throughout.
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));

template <int P> struct ThetaArith;
template <> struct ThetaArith<0> {
  using R = double; using Acc = v4d;
  static __device__ __forceinline__ int drow(int lane, int e) { return (lane >> 4) + 4 * e; }
};
template <> struct ThetaArith<1> {
  using R = float; using Acc = v4f;
  static __device__ __forceinline__ int drow(int lane, int e) { return 4 * (lane >> 4) + e; }
};
template <> struct ThetaArith<2> {
  using R = float; using Acc = v4f;
  static __device__ __forceinline__ int drow(int lane, int e) { return 4 * (lane >> 4) + e; }
};

// BF16 of a float, rounded to nearest even (finite inputs), as raw bits.
__device__ __forceinline__ short bf16_bits(float x) {
  const unsigned u = __float_as_uint(x);
  return (short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_value(short h) {
  return __uint_as_float(((unsigned)(unsigned short)h) << 16);
}
// x = hi + lo to about 2^-17 relative
__device__ __forceinline__ void bf16_split(float x, short& hi, short& lo) {
  hi = bf16_bits(x);
  lo = bf16_bits(x - bf16_value(hi));
}

constexpr int kThetaKC = 16;  // K values per LDS stage

template <int TM>
struct ThetaSmemDims {
  // staging: Ar, Ai [TM][KC+1]; Br, Bi [KC][TM+1]; epilogue: Or, Oi [TM][TM+1]
  static constexpr int stage = 2 * TM * (kThetaKC + 1) + 2 * kThetaKC * (TM + 1);
  static constexpr int out = 2 * TM * (TM + 1);
  static constexpr int elems = stage > out ? stage : out;
};

template <int P, int TM>
__device__ void theta_tiles(const UpdDesc& d, int chiL, int chiM, int chiR, int orient,
                            double* smraw) {
  using Ar = ThetaArith<P>;
  using R = typename Ar::R;
  using Acc = typename Ar::Acc;
  constexpr int KC = kThetaKC;
  constexpr int TV = TM / 2;   // v (and w) values per tile
  constexpr int NS = TM / 32;  // 16 x 16 subtiles per wave edge
  R* sar = reinterpret_cast<R*>(smraw);
  R* sai = sar + TM * (KC + 1);
  R* sbr = sai + TM * (KC + 1);
  R* sbi = sbr + KC * (TM + 1);
  R* sor = reinterpret_cast<R*>(smraw);
  R* soi = sor + TM * (TM + 1);
  auto AR = [&](int r, int k) -> R& { return sar[r * (KC + 1) + k]; };
  auto AI = [&](int r, int k) -> R& { return sai[r * (KC + 1) + k]; };
  auto BR = [&](int k, int c) -> R& { return sbr[k * (TM + 1) + c]; };
  auto BI = [&](int k, int c) -> R& { return sbi[k * (TM + 1) + c]; };
  auto OR = [&](int r, int c) -> R& { return sor[r * (TM + 1) + c]; };
  auto OI = [&](int r, int c) -> R& { return soi[r * (TM + 1) + c]; };

  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wr = (wid >> 1) * (TM / 2), wc = (wid & 1) * (TM / 2);
  const int nvt = (chiL + TV - 1) / TV, nwt = (chiR + TV - 1) / TV;
  const double2* g = d.gsrc ? d.gsrc : d.gate;
  for (int tile = blockIdx.x; tile < nvt * nwt; tile += gridDim.x) {
    const int v0 = (tile / nwt) * TV, w0 = (tile % nwt) * TV;
    Acc cre[NS][NS], cim[NS][NS];
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
      for (int j = 0; j < NS; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) cre[i][j][e] = cim[i][j][e] = (R)0;
    for (int k0 = 0; k0 < chiM; k0 += KC) {
      __syncthreads();  // the previous stage (or tile epilogue) is read
      for (int e = tid; e < TM * KC; e += 256) {
        const int r = e / KC, kk = e % KC;
        const int row = 2 * v0 + r, u = k0 + kk;
        double2 x = make_double2(0.0, 0.0);
        if (row < 2 * chiL && u < chiM) x = d.gk[(size_t)row * chiM + u];
        AR(r, kk) = (R)x.x;
        AI(r, kk) = (R)x.y;
      }
      for (int e = tid; e < KC * TM; e += 256) {
        const int kk = e / TM, c = e % TM;
        const int t = c / TV, w = w0 + c % TV, u = k0 + kk;
        double2 x = make_double2(0.0, 0.0);
        if (u < chiM && w < chiR) x = cscale(d.gk1[((size_t)u * 2 + t) * chiR + w], d.lamM[u]);
        BR(kk, c) = (R)x.x;
        BI(kk, c) = (R)x.y;
      }
      __syncthreads();
      if constexpr (P == 2) {
        v4s arh[NS], arl[NS], aih[NS], ail[NS], anh[NS], anl[NS];
#pragma unroll
        for (int i = 0; i < NS; ++i) {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            short h, l;
            bf16_split(AR(wr + 16 * i + li, 4 * lk + e), h, l);
            arh[i][e] = h; arl[i][e] = l;
            bf16_split(AI(wr + 16 * i + li, 4 * lk + e), h, l);
            aih[i][e] = h; ail[i][e] = l;
            anh[i][e] = (short)(h ^ (short)0x8000);  // -Ai: sign bit flipped
            anl[i][e] = (short)(l ^ (short)0x8000);
          }
        }
#pragma unroll
        for (int j = 0; j < NS; ++j) {
          v4s brh, brl, bih, bil;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            short h, l;
            bf16_split(BR(4 * lk + e, wc + 16 * j + li), h, l);
            brh[e] = h; brl[e] = l;
            bf16_split(BI(4 * lk + e, wc + 16 * j + li), h, l);
            bih[e] = h; bil[e] = l;
          }
#pragma unroll
          for (int i = 0; i < NS; ++i) {
            // small terms first
            cre[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arl[i], brh, cre[i][j], 0, 0, 0);
            cre[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh[i], brl, cre[i][j], 0, 0, 0);
            cre[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(anl[i], bih, cre[i][j], 0, 0, 0);
            cre[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(anh[i], bil, cre[i][j], 0, 0, 0);
            cre[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh[i], brh, cre[i][j], 0, 0, 0);
            cre[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(anh[i], bih, cre[i][j], 0, 0, 0);
            cim[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arl[i], bih, cim[i][j], 0, 0, 0);
            cim[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh[i], bil, cim[i][j], 0, 0, 0);
            cim[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ail[i], brh, cim[i][j], 0, 0, 0);
            cim[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aih[i], brl, cim[i][j], 0, 0, 0);
            cim[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh[i], bih, cim[i][j], 0, 0, 0);
            cim[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aih[i], brh, cim[i][j], 0, 0, 0);
          }
        }
      } else {
#pragma unroll
        for (int k4 = 0; k4 < KC / 4; ++k4) {
          R a_r[NS], a_i[NS], b_r[NS], b_i[NS];
#pragma unroll
          for (int i = 0; i < NS; ++i) {
            a_r[i] = AR(wr + 16 * i + li, 4 * k4 + lk);
            a_i[i] = AI(wr + 16 * i + li, 4 * k4 + lk);
          }
#pragma unroll
          for (int j = 0; j < NS; ++j) {
            b_r[j] = BR(4 * k4 + lk, wc + 16 * j + li);
            b_i[j] = BI(4 * k4 + lk, wc + 16 * j + li);
          }
#pragma unroll
          for (int i = 0; i < NS; ++i)
#pragma unroll
            for (int j = 0; j < NS; ++j) {
              if constexpr (P == 0) {
                cre[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r[i], b_r[j], cre[i][j], 0, 0, 0);
                cre[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a_i[i], b_i[j], cre[i][j], 0, 0, 0);
                cim[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_r[i], b_i[j], cim[i][j], 0, 0, 0);
                cim[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_i[i], b_r[j], cim[i][j], 0, 0, 0);
              } else {
                cre[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_r[i], b_r[j], cre[i][j], 0, 0, 0);
                cre[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(-a_i[i], b_i[j], cre[i][j], 0, 0, 0);
                cim[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_r[i], b_i[j], cim[i][j], 0, 0, 0);
                cim[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_i[i], b_r[j], cim[i][j], 0, 0, 0);
              }
            }
        }
      }
    }
    __syncthreads();  // the last stage is read before the tile overwrites it
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
      for (int j = 0; j < NS; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int r = wr + 16 * i + Ar::drow(lane, e), c = wc + 16 * j + li;
          OR(r, c) = cre[i][j][e];
          OI(r, c) = cim[i][j][e];
        }
    __syncthreads();
    // epilogue per (v, w): lamL, lamR, the gate, the store into X.  Threads
    // run along v for theta (X column-major: rows 2v, 2v+1 of a column are
    // adjacent) and along w for theta^H.
    for (int p = tid; p < TV * TV; p += 256) {
      const int vv = orient == 0 ? p % TV : p / TV;
      const int ww = orient == 0 ? p / TV : p % TV;
      const int v = v0 + vv, w = w0 + ww;
      if (v >= chiL || w >= chiR) continue;
      R ax[4], ay[4];
#pragma unroll
      for (int st = 0; st < 4; ++st) {
        const int s = st >> 1, t = st & 1;
        ax[st] = OR(2 * vv + s, t * TV + ww);
        ay[st] = OI(2 * vv + s, t * TV + ww);
      }
      R f = (R)1;
      if (d.lamL) f *= (R)d.lamL[v];
      if (d.lamR) f *= (R)d.lamR[w];
#pragma unroll
      for (int st = 0; st < 4; ++st) { ax[st] *= f; ay[st] *= f; }
      double2 out[4];
      if (d.has_gate) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          R ox = (R)0, oy = (R)0;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const R gx = (R)g[i * 4 + j].x, gy = (R)g[i * 4 + j].y;
            ox += gx * ax[j] - gy * ay[j];
            oy += gx * ay[j] + gy * ax[j];
          }
          out[i] = make_double2((double)ox, (double)oy);
        }
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) out[i] = make_double2((double)ax[i], (double)ay[i]);
      }
#pragma unroll
      for (int st = 0; st < 4; ++st) {
        const int s = st >> 1, t = st & 1;
        const int row = v * 2 + s, col = t * chiR + w;
        if (orient == 0)
          d.X[(size_t)col * d.ldx + row] = out[st];
        else
          d.X[(size_t)row * d.ldx + col] = cconj(out[st]);
      }
    }
  }
}

// theta'[v,(s,t),w] = sum_{s't'} U[(s,t),(s',t')] lamL_v lamR_w
//                     sum_u Gk[v,s',u] lamM_u Gk1[u,t',w]
// written into X as theta (orient 0, m >= n) or theta^H (orient 1), on the
// matrix cores in TM x TM tiles (TM = 32 or 64).  In the factored modes X is
// filled elementwise instead (no contraction):
//   mode 1: X[(t,w), u] = conj(lamM_u Gk1[u,t,w] lamR_w)
//   mode 2: X[(v,s), u] = lamL_v Gk[v,s,u] lamM_u
template <int TM>
__global__ __launch_bounds__(256) void k_theta(const UpdDesc* __restrict__ descs, int nt) {
  const UpdDesc& d = descs[blockIdx.y];
  const int chiL = d.dims[0], chiM = d.dims[1], chiR = d.dims[2];
  const TaskShape ts = task_shape(d);
  const int orient = ts.orient;
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) {
      d.info[kInfoM] = ts.m;
      d.info[kInfoN] = ts.n;
      d.info[kInfoR] = ts.R;
      d.info[kInfoRJ] = ts.R;
      d.info[kInfoNorms] = 0;
      d.info[kInfoC] = ts.C;
      d.info[kInfoOrient] = orient;
      d.info[kInfoMode] = ts.mode;
      d.info[kInfoFlags] = 0;
      d.info[kInfoSweeps] = 0;
      d.info[kInfoChiL] = chiL;
      d.info[kInfoChiM] = chiM;
      d.info[kInfoChiR] = chiR;
    }
    for (int s = threadIdx.x; s < kMaxSweeps; s += blockDim.x) d.offs[s * nt] = 0.0;
  }
  if (ts.mode != 0) {
    // element counts stay below 2^32 (R <= 2048 rows, chiM <= 1024): 32-bit
    // index arithmetic
    const unsigned tot = (unsigned)ts.R * (unsigned)chiM;
    const unsigned stride = gridDim.x * blockDim.x;
    for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += stride) {
      if (ts.mode == 1) {
        const int u = (int)(e / (unsigned)ts.n), r = (int)(e % (unsigned)ts.n);
        double f = d.lamM[u];
        if (d.lamR) f *= d.lamR[r % chiR];
        d.X[(size_t)u * d.ldx + r] = cconj(cscale(d.gk1[e], f));
      } else {
        // e runs down X's columns, so the writes are coalesced; Gamma[k]
        // (row-major, 2 MiB at chi = 256) is read across its rows from L2
        const int u = (int)(e / (unsigned)ts.m), r = (int)(e % (unsigned)ts.m);
        double f = d.lamM[u];
        if (d.lamL) f *= d.lamL[r >> 1];
        d.X[(size_t)u * d.ldx + r] = cscale(d.gk[(size_t)r * chiM + u], f);
      }
    }
    return;
  }
  __shared__ double smraw[ThetaSmemDims<TM>::elems];
  if (d.prec == 2)
    theta_tiles<2, TM>(d, chiL, chiM, chiR, orient, smraw);
  else if (d.prec == 1)
    theta_tiles<1, TM>(d, chiL, chiM, chiR, orient, smraw);
  else
    theta_tiles<0, TM>(d, chiL, chiM, chiR, orient, smraw);
}

// ---------------------------------------------------------------- W = I
__global__ __launch_bounds__(256) void k_winit(const UpdDesc* __restrict__ descs) {
  const UpdDesc& d = descs[blockIdx.y];
  const int C = task_shape(d).C;
  const unsigned tot = (unsigned)C * (unsigned)C;  // C <= 1024
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += gridDim.x * blockDim.x) {
    const int i = (int)(e % (unsigned)C), j = (int)(e / (unsigned)C);
    d.W[(size_t)j * d.ldw + i] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
}

// ------------------------------------------------------------- finish
constexpr double kF32NoiseFloor = 9.5367431640625e-07;  // 2^-20
constexpr double kBF16SplitNoiseFloor = 1.220703125e-04;  // 2^-13

// Column norms of X are the singular values (theta = U S V^H).  Sort them
// descending, apply process_svd (mps.rs:236-293): normalise, pick the rank,
// renormalise, and record the right-factor scale fac_j = s_j * norm2.
__global__ __launch_bounds__(1024) void k_finish(const UpdDesc* __restrict__ descs, int nt,
                                                 int sweeps_run) {
  const UpdDesc& d = descs[blockIdx.x];
  const int R = d.info[kInfoR], C = d.info[kInfoC];
  __shared__ double key[kMaxCols];
  __shared__ int idx[kMaxCols];
  __shared__ double red[16];
  __shared__ int first_bad;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;

  if (tid == 0) d.info[kInfoSweeps] = sweeps_run;

  // the column norms, unless the QR preconditioner's last kernel has formed
  // them by this arithmetic (bitwise the same) while X was in its cache
  if (d.info[kInfoNorms] == 0) {
    for (int j = wid; j < C; j += 16) {
      double s = 0.0;
      for (int r = lane; r < R; r += 64) s += cabs2(d.X[(size_t)j * d.ldx + r]);
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
      if (lane == 0) d.sraw[j] = sqrt(s);
    }
  }
  __syncthreads();
  __threadfence_block();
  int P2 = 1;
  while (P2 < C) P2 <<= 1;
  for (int i = tid; i < P2; i += blockDim.x) {
    key[i] = (i < C) ? d.sraw[i] : -1.0;
    idx[i] = i;
  }
  __syncthreads();
  // descending key, ascending index on ties
  bitonic_desc(key, idx, P2);
  // norm of all values (mps.rs:243-248)
  double v = 0.0;
  for (int i = tid; i < C; i += blockDim.x) v += key[i] * key[i];
  const double norm = sqrt(block_sum(v, red));
  for (int i = tid; i < C; i += blockDim.x) key[i] = key[i] / norm;
  if (tid == 0) first_bad = C;
  __syncthreads();
  // leading count of normal values above the threshold (take_while)
  double thr;
  if (d.tkind == 2) thr = fabs(d.teps);
  else thr = DBL_EPSILON;
  // With an FP32 contraction, normalised singular values below the FP32
  // round-off of theta (2^-20 relative) are noise, not Schmidt values:
  // they are not counted as kept values of the rank rule.  The BF16-split
  // contraction keeps about 16 bits per product: its floor is 2^-13.
  if (d.prec == 1) thr = fmax(thr, kF32NoiseFloor);
  if (d.prec == 2) thr = fmax(thr, kBF16SplitNoiseFloor);
  for (int i = tid; i < C; i += blockDim.x) {
    const double s = key[i];
    const bool ok = (s >= DBL_MIN) && isfinite(s) && (s > thr);
    if (!ok) atomicMin(&first_bad, i);
  }
  __syncthreads();
  const int lead = first_bad;
  int rank;
  if (d.tkind == 1) {
    const int r = d.tchi < 1 ? 1 : d.tchi;
    rank = r < lead ? r : lead;
  } else {
    rank = lead > 1 ? lead : 1;
  }
  int flags = 0;
  if (rank > d.cap_rank) {
    rank = d.cap_rank;
    flags |= kFlagCapacity;
  }
  if (rank > C) rank = C;
  double v2 = 0.0;
  for (int i = tid; i < rank; i += blockDim.x) v2 += key[i] * key[i];
  const double norm2 = sqrt(block_sum(v2, red));
  for (int i = tid; i < rank; i += blockDim.x) {
    const double s = key[i] / norm2;
    d.sval[i] = s;
    d.lamM[i] = s;
    d.fac[i] = s * norm2;
    d.perm[i] = idx[i];
  }
  if (tid == 0) {
    d.info[kInfoRank] = rank;
    d.info[kInfoFlags] |= flags;
    d.dims[1] = rank;
    // a clamped rank is reported through the sticky words the host reads
    // when the call that enqueued this update returns
    if (flags) {
      atomicOr(d.hflags, flags);
      atomicOr(d.eflags, flags);
    }
  }
}

// ------------------------------------------------------ factor product
// Factored modes, after k_finish: the kept singular vectors times the
// isometric factor, for the kept columns j < rank.
//   mode 1: T[j][row] = sum_u Gk[row, u] W[u, perm_j]          (row < m)
//           = Gamma[k] of lamL Gamma[k] U_M (already divided by lamL)
//   mode 2: T[j][c]   = sum_u conj(W[u, perm_j]) Gk1[u, c]     (c < n)
//           = rows of V_A^H Gamma[k+1] (already divided by lamR)
// Tile: 64 values of the free index x 4 JPT values of j per 256-thread
// block (JPT per thread), u streamed through LDS 16 at a time.  Grid (x
// tiles, j tiles, task).  Every output sums over u in the same order for
// any JPT: launches of few tasks take JPT = 2 (four times the workgroups of
// JPT = 8 for a lone update's product), batches JPT = 8 (a quarter of the
// reloads of the factor).
template <int JPT>
__global__ __launch_bounds__(256) void k_fmul(const UpdDesc* __restrict__ descs) {
  constexpr int JT = 4 * JPT;  // values of j per block
  const UpdDesc& d = descs[blockIdx.z];
  const int mode = d.info[kInfoMode];
  if (mode == 0) return;
  const int K = d.info[kInfoChiM];
  const int rank = d.info[kInfoRank];
  const int nx = (mode == 1) ? d.info[kInfoM] : d.info[kInfoN];
  const int x0 = blockIdx.x * 64, j0 = blockIdx.y * JT;
  if (x0 >= nx || j0 >= rank) return;
  __shared__ double2 Ps[16][65];      // [u][x]
  __shared__ double2 Ws[16][JT + 1];  // [u][j]
  __shared__ int pj[JT];
  const int tid = threadIdx.x;
  if (tid < JT) pj[tid] = (j0 + tid < rank) ? d.perm[j0 + tid] : -1;
  __syncthreads();
  const int tx = tid & 63, ty = tid >> 6;
  double2 acc[JPT];
#pragma unroll
  for (int k = 0; k < JPT; ++k) acc[k] = make_double2(0.0, 0.0);
  const double2* __restrict__ G = (mode == 1) ? d.gk : d.gk1;
  for (int u0 = 0; u0 < K; u0 += 16) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      int uu, xx;
      if (mode == 1) {  // Gk[x * K + u]: u contiguous
        uu = e & 15;
        xx = e >> 4;
      } else {          // Gk1[u * n + x]: x contiguous
        xx = e & 63;
        uu = e >> 6;
      }
      const int u = u0 + uu, x = x0 + xx;
      double2 v = make_double2(0.0, 0.0);
      if (u < K && x < nx) v = (mode == 1) ? G[(size_t)x * K + u] : G[(size_t)u * nx + x];
      Ps[uu][xx] = v;
    }
#pragma unroll
    for (int q = 0; q < (16 * JT + 255) / 256; ++q) {
      const int e = tid + 256 * q;
      if (e < 16 * JT) {
        const int uu = e & 15, jj = e >> 4;
        const int u = u0 + uu, c = pj[jj];
        double2 w = make_double2(0.0, 0.0);
        if (c >= 0 && u < K) w = d.W[(size_t)c * d.ldw + u];
        Ws[uu][jj] = (mode == 2) ? cconj(w) : w;
      }
    }
    __syncthreads();
#pragma unroll 4
    for (int uu = 0; uu < 16; ++uu) {
      const double2 pv = Ps[uu][tx];
#pragma unroll
      for (int k = 0; k < JPT; ++k) acc[k] = cadd(acc[k], cmul(pv, Ws[uu][ty * JPT + k]));
    }
    __syncthreads();
  }
  const int x = x0 + tx;
  if (x < nx) {
#pragma unroll
    for (int k = 0; k < JPT; ++k) {
      const int j = j0 + ty * JPT + k;
      if (j < rank) d.T[(size_t)j * d.ldx + x] = acc[k];
    }
  }
}

// -------------------------------------------------------------- split
// Gamma[k][v,s,j]   = U[(v,s), j] / lamL_v
// Gamma[k+1][j,t,w] = (fac_j V^H[j, (t,w)]) / s_j / lamR_w
// plus the partial sums of local_norm for both sites.  U and V^H come from
//   full mode, orient 0: U = X / sraw, V = W;  orient 1: U = W, V = X / sraw
//   mode 1: Gamma[k] = T (lamL already divided out), V = X / sraw
//   mode 2: U = X / sraw, V^H / lamR = T
__global__ __launch_bounds__(256) void k_split(const UpdDesc* __restrict__ descs) {
  const UpdDesc& d = descs[blockIdx.y];
  const int chiL = d.info[kInfoChiL], chiR = d.info[kInfoChiR];
  const int orient = d.info[kInfoOrient];
  const int mode = d.info[kInfoMode];
  const int rank = d.info[kInfoRank];
  __shared__ double red[4];
  // element counts stay below 2^32 (chi <= 1024, rank <= 1024): 32-bit
  // index arithmetic
  const unsigned nk = (unsigned)chiL * 2u * (unsigned)rank;
  const unsigned nk1 = (unsigned)rank * 2u * (unsigned)chiR;
  const unsigned stride = gridDim.x * blockDim.x;
  double accA = 0.0, accB = 0.0;
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < nk; e += stride) {
    const int j = (int)(e % (unsigned)rank);
    const int row = (int)(e / (unsigned)rank);  // v*2 + s
    const int v = row >> 1;
    const int pj = d.perm[j];
    const double wl = d.lamL ? d.lamL[v] : 1.0;
    double2 u;
    if (mode == 1) {
      u = d.T[(size_t)j * d.ldx + row];
    } else {
      if (orient == 0) {
        u = cdivr(d.X[(size_t)pj * d.ldx + row], d.sraw[pj]);
      } else {
        u = d.W[(size_t)pj * d.ldw + row];
      }
      if (d.lamL) u = cdivr(u, wl);
    }
    d.gk[e] = u;
    const double sj = d.sval[j];
    accA += cabs2(u) * (wl * wl) * (sj * sj);
  }
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < nk1; e += stride) {
    const int j = (int)(e / (2u * (unsigned)chiR));
    const int cidx = (int)(e % (2u * (unsigned)chiR));  // t*chiR + w
    const int w = cidx < chiR ? cidx : cidx - chiR;
    const int pj = d.perm[j];
    const double sj = d.sval[j];
    const double wr = d.lamR ? d.lamR[w] : 1.0;
    double2 g;
    if (mode == 2) {
      g = cdivr(cscale(d.T[(size_t)j * d.ldx + cidx], d.fac[j]), sj);
    } else {
      double2 vh;
      if (orient == 0) {
        vh = cconj(d.W[(size_t)pj * d.ldw + cidx]);
      } else {
        vh = cconj(cdivr(d.X[(size_t)pj * d.ldx + cidx], d.sraw[pj]));
      }
      g = cdivr(cscale(vh, d.fac[j]), sj);
      if (d.lamR) g = cdivr(g, wr);
    }
    d.gk1[e] = g;
    accB += cabs2(g) * (sj * sj) * (wr * wr);
  }
  const double sA = block_sum(accA, red);
  const double sB = block_sum(accB, red);
  if (threadIdx.x == 0) {
    d.part[2 * blockIdx.x + 0] = sA;
    d.part[2 * blockIdx.x + 1] = sB;
  }
}

__global__ __launch_bounds__(256) void k_renorm(const UpdDesc* __restrict__ descs) {
  const UpdDesc& d = descs[blockIdx.y];
  const int chiL = d.info[kInfoChiL], chiR = d.info[kInfoChiR];
  const int rank = d.info[kInfoRank];
  __shared__ double nrm[2];
  if (threadIdx.x < 64) {
    // the split kernel's per-block partial norms, one block per lane, summed
    // across the wave in a fixed order
    double a = 0.0, b = 0.0;
    for (int q = threadIdx.x; q < (int)gridDim.x; q += 64) {
      a += d.part[2 * q + 0];
      b += d.part[2 * q + 1];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_xor(a, off, 64);
      b += __shfl_xor(b, off, 64);
    }
    if (threadIdx.x == 0) {
      nrm[0] = sqrt(a);
      nrm[1] = sqrt(b);
    }
  }
  __syncthreads();
  const double nA = nrm[0], nB = nrm[1];
  const unsigned nk = (unsigned)chiL * 2u * (unsigned)rank;
  const unsigned nk1 = (unsigned)rank * 2u * (unsigned)chiR;
  const unsigned stride = gridDim.x * blockDim.x;
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < nk; e += stride)
    d.gk[e] = cdivr(d.gk[e], nA);
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < nk1; e += stride)
    d.gk1[e] = cdivr(d.gk1[e], nB);
}

// ------------------------------------------------------------ measure
// A site's Born probabilities and projection (mps.rs:1027-1134, 1427-1465)
// over gridDim.x workgroups per site (grid x, site y): k_measure_part sums
// p_s = sum_v lamL_v^2 sum_u |Gamma[v, s, u]|^2 lamR_u^2 over the workgroup's
// rows v; k_measure_project sums the partials in one fixed order (so every
// workgroup of the site holds bitwise the same p0, p1 and outcome), the
// first one writes the results, and each projects its own rows.  A wave
// takes whole rows, its lanes run along u, so (v, 0) and (v, 1) are read
// contiguously.
__device__ __forceinline__ void meas_rows(int chiL, int g, int G, int& v0, int& v1) {
  const int per = (chiL + G - 1) / G;
  v0 = min(chiL, g * per);
  v1 = min(chiL, v0 + per);
}

__global__ __launch_bounds__(256) void k_measure_part(const MeasDesc* __restrict__ descs,
                                                      double* __restrict__ parts) {
  const MeasDesc& d = descs[blockIdx.y];
  const int chiL = d.dims[0], chiR = d.dims[1];
  __shared__ double red[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int v0, v1;
  meas_rows(chiL, blockIdx.x, gridDim.x, v0, v1);
  double p0 = 0.0, p1 = 0.0;
  for (int v = v0 + wid; v < v1; v += 4) {
    const double wl = d.lamL ? d.lamL[v] * d.lamL[v] : 1.0;
    const double2* r0 = d.g + (size_t)v * 2 * chiR;
    const double2* r1 = r0 + chiR;
    double q0 = 0.0, q1 = 0.0;
    for (int u = lane; u < chiR; u += 64) {
      const double wr = d.lamR ? d.lamR[u] * d.lamR[u] : 1.0;
      q0 += cabs2(r0[u]) * wr;
      q1 += cabs2(r1[u]) * wr;
    }
    p0 += q0 * wl;
    p1 += q1 * wl;
  }
  p0 = block_sum(p0, red);
  p1 = block_sum(p1, red);
  if (threadIdx.x == 0) {
    double* pp = parts + ((size_t)blockIdx.y * kMeasParts + blockIdx.x) * 2;
    pp[0] = p0;
    pp[1] = p1;
  }
}

__global__ __launch_bounds__(256) void k_measure_project(const MeasDesc* __restrict__ descs,
                                                         const double* __restrict__ parts) {
  const MeasDesc& d = descs[blockIdx.y];
  const int chiL = d.dims[0], chiR = d.dims[1];
  __shared__ double s_p[2];
  if (threadIdx.x < 64) {
    // the site's partials, one per lane, summed across the wave in a fixed order
    const double* pp = parts + (size_t)blockIdx.y * kMeasParts * 2;
    double a = 0.0, b = 0.0;
    if (threadIdx.x < (int)gridDim.x) {
      a = pp[2 * threadIdx.x];
      b = pp[2 * threadIdx.x + 1];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_xor(a, off, 64);
      b += __shfl_xor(b, off, 64);
    }
    if (threadIdx.x == 0) {
      s_p[0] = a;
      s_p[1] = b;
    }
  }
  __syncthreads();
  const double p0 = s_p[0], p1 = s_p[1];
  int flag = 0, out;
  double prob;
  if (d.forced >= 0) {
    out = d.forced;
    prob = out == 0 ? p0 : p1;
  } else {
    // first j with r < cumulative probability (mps.rs:1441-1446)
    const double c0 = p0, c1 = p0 + p1;
    if (d.r < c0) out = 0;
    else if (d.r < c1) out = 1;
    else {
      out = p1 > 0.0 ? 1 : 0;
      flag |= 2;
    }
    prob = out == 0 ? p0 : p1;
  }
  if (d.forced != -2 && !(prob > 0.0)) flag |= 1;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    d.out_outcome[d.slot] = out;
    d.out_prob[d.slot] = prob;
    d.out_probs[2 * d.slot + 0] = p0;
    d.out_probs[2 * d.slot + 1] = p1;
    d.out_flag[d.slot] = flag;
    if (d.out_dims) {
      d.out_dims[2 * blockIdx.y + 0] = chiL;
      d.out_dims[2 * blockIdx.y + 1] = chiR;
    }
  }
  if (d.forced == -2) return;
  const double renorm = sqrt(prob);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int v0, v1;
  meas_rows(chiL, blockIdx.x, gridDim.x, v0, v1);
  for (int v = v0 + wid; v < v1; v += 4) {
    double2* keep = d.g + ((size_t)v * 2 + out) * chiR;
    double2* drop = d.g + ((size_t)v * 2 + (1 - out)) * chiR;
    for (int u = lane; u < chiR; u += 64) {
      keep[u] = cdivr(keep[u], renorm);
      drop[u] = make_double2(0.0, 0.0);
    }
  }
}

// ------------------------------------------------------------ entropy
// One workgroup per read-out: a == 1: -sum s^2 ln s^2; otherwise
// ln(sum (s^2)^a) / (1 - a).
__global__ __launch_bounds__(256) void k_entropy(const EntDesc* __restrict__ descs,
                                                 double* out) {
  __shared__ double red[4];
  const EntDesc q = descs[blockIdx.x];
  const int len = *q.dim;
  double acc = 0.0;
  for (int i = threadIdx.x; i < len; i += blockDim.x) {
    const double s2 = q.lam[i] * q.lam[i];
    if (q.a == 1.0) acc += -s2 * log(s2);
    else acc += pow(s2, q.a);
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) out[q.slot] = (q.a == 1.0) ? s : log(s) * (1.0 / (1.0 - q.a));
}

// ------------------------------------------------------------ contract
// One site of MPS::contract (mps.rs:794-798): out[(p, s), r] =
// sum_l acc[p, l] lamL[l] Gamma[l, s, r] for p < P, with chiL = dims[0] and
// chiR = dims[1]; acc holds P x chiL, out 2P x chiR (qubit 0 slowest).
__global__ __launch_bounds__(256) void k_contract(const double2* __restrict__ acc, size_t P,
                                                  const int* __restrict__ dims,
                                                  const double2* __restrict__ g,
                                                  const double* __restrict__ lamL,
                                                  double2* __restrict__ out) {
  const int chiL = dims[0], chiR = dims[1];
  const size_t tot = 2 * P * (size_t)chiR;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t ps = e / chiR;
    const int r = (int)(e % chiR);
    const size_t p = ps >> 1;
    const int s = (int)(ps & 1);
    double2 o = make_double2(0.0, 0.0);
    for (int l = 0; l < chiL; ++l) {
      double2 a = acc[p * chiL + l];
      if (lamL) a = cscale(a, lamL[l]);
      o = cadd(o, cmul(a, g[((size_t)l * 2 + s) * chiR + r]));
    }
    out[e] = o;
  }
}

// ------------------------------------------------------ matrix part
// rho[i][j] = sum_(l, r) w[r] T[l][i][r] conj(T[l][j][r]) over T of L x D x R
// (row-major), w[r] = lamR[r]^2 (1 without lamR): the reduced density matrix
// of D = 2^(b-a) states of sites a..b-1 from T = lambda[a-1] Gamma[a] ...
// Gamma[b-1] and the right boundary's lambda.  16 x 16 tiles of rho per
// workgroup, the summed index (l, r) staged through LDS 16 at a time.
__global__ __launch_bounds__(256) void k_rho_part(const double2* __restrict__ t, size_t D,
                                                  size_t R, size_t E,
                                                  const double* __restrict__ lamR,
                                                  double2* __restrict__ rho) {
  __shared__ double2 mi[16][17], mj[16][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const size_t i0 = (size_t)blockIdx.y * 16, j0 = (size_t)blockIdx.x * 16;
  double2 acc = make_double2(0.0, 0.0);
  for (size_t e0 = 0; e0 < E; e0 += 16) {
    // thread (ty, tx) stages w T[i0 + ty][e0 + tx] and T[j0 + ty][e0 + tx]
    const size_t e = e0 + tx;
    const size_t l = e / R, r = e % R;
    const size_t ii = i0 + ty, jj = j0 + ty;
    const double w = (lamR && e < E) ? lamR[r] * lamR[r] : 1.0;
    mi[ty][tx] = (ii < D && e < E) ? cscale(t[(l * D + ii) * R + r], w) : make_double2(0.0, 0.0);
    mj[ty][tx] = (jj < D && e < E) ? t[(l * D + jj) * R + r] : make_double2(0.0, 0.0);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) acc = cadd(acc, cmulc(mj[tx][k], mi[ty][k]));
    __syncthreads();
  }
  const size_t i = i0 + ty, j = j0 + tx;
  if (i < D && j < D) rho[i * D + j] = acc;
}

// ------------------------------------------------------ from_vector
// The SVD task of one factorization step: info (R, C, orientation), zeroed
// per-sweep maxima, W = I; in orientation B also X = q, the transpose of
// the conjugated factor in memory.
__global__ __launch_bounds__(256) void k_fv_init(FvArgs a) {
  const unsigned tid = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned stride = gridDim.x * blockDim.x;
  if (tid == 0) {
    for (int i = 0; i < kInfoLen; ++i) a.info[i] = 0;
    a.info[kInfoM] = a.m;
    a.info[kInfoN] = a.ncols;
    a.info[kInfoR] = a.R;
    a.info[kInfoRJ] = a.R;
    a.info[kInfoC] = a.C;
    a.info[kInfoOrient] = a.orient;
  }
  for (unsigned s = tid; s < (unsigned)kMaxSweeps; s += stride) a.offs[s] = 0.0;
  const unsigned cc = (unsigned)a.C * (unsigned)a.C;
  for (unsigned e = tid; e < cc; e += stride) {
    const unsigned i = e % (unsigned)a.C, j = e / (unsigned)a.C;
    a.W[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  if (a.orient == 1) {
    // X[c][i] (column-major, ld m) = q[i][c] = conj(src[i * ncols + c])
    const size_t tot = (size_t)a.m * a.ncols;
    for (size_t e = tid; e < tot; e += stride) {
      const size_t i = e % (size_t)a.m, c = e / (size_t)a.m;
      a.X[e] = cconj(a.src[i * a.ncols + c]);
    }
  }
}

// Gamma[k][v, s, j] = U[(2v + s), perm_j] / lamPrev_v and the next factor
// conj(q_new)[j][c] with q_new[j] = fac_j V^H[perm_j] (process_svd's row
// scale), for the kept j < rank.  Orientation A (X = q^H): U = W,
// V = X W / sraw.  Orientation B (X = q): U = X W / sraw, V = W.
__global__ __launch_bounds__(256) void k_fv_split(FvArgs a) {
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const int r = a.rank;
  const size_t ng = (size_t)a.m * r;
  for (size_t e = tid; e < ng; e += stride) {
    const int j = (int)(e % (size_t)r), i = (int)(e / (size_t)r);  // i = 2v + s
    const int pj = a.perm[j];
    double2 u = a.orient == 0 ? a.W[(size_t)pj * a.C + i]
                              : cdivr(a.X[(size_t)pj * a.R + i], a.sraw[pj]);
    if (a.lamPrev) u = cdivr(u, a.lamPrev[i >> 1]);
    a.gk[e] = u;
  }
  const size_t nq = (size_t)r * a.ncols;
  for (size_t e = tid; e < nq; e += stride) {
    const int c = (int)(e % (size_t)a.ncols), j = (int)(e / (size_t)a.ncols);
    const int pj = a.perm[j];
    double2 q;
    if (a.orient == 0) {
      // conj(q_new[j][c]) = fac_j (X W)[c][pj] / sraw_pj
      q = cscale(cdivr(a.X[(size_t)pj * a.R + c], a.sraw[pj]), a.fac[j]);
    } else {
      // conj(q_new[j][c]) = fac_j W[c][pj]
      q = cscale(a.W[(size_t)pj * a.C + c], a.fac[j]);
    }
    a.X2[e] = q;
  }
}

// Gamma[n-1][v, s, 0] = q[v][s] / lam_v from the conjugated final factor.
__global__ __launch_bounds__(256) void k_fv_last(const double2* __restrict__ q,
                                                 const double* __restrict__ lam, int udim,
                                                 double2* __restrict__ g) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < 2 * udim) g[e] = cdivr(cconj(q[e]), lam[e >> 1]);
}

// ----------------------------------------------------------------- u1
struct Gate2 { double2 m[4]; };
__global__ __launch_bounds__(256) void k_u1(double2* g, const int* dims, Gate2 op) {
  const int chiL = dims[0], chiR = dims[1];
  const size_t tot = (size_t)chiL * chiR;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int v = (int)(e / chiR), u = (int)(e % chiR);
    double2* p0 = g + ((size_t)v * 2 + 0) * chiR + u;
    double2* p1 = g + ((size_t)v * 2 + 1) * chiR + u;
    const double2 a = *p0, b = *p1;
    *p0 = cadd(cmul(op.m[0], a), cmul(op.m[1], b));
    *p1 = cadd(cmul(op.m[2], a), cmul(op.m[3], b));
  }
}

// ------------------------------------------------------------- launch
void launch_theta(const UpdDesc* descs, int nt, int tm, int gx, hipStream_t s) {
  if (tm == 32)
    hipLaunchKernelGGL(k_theta<32>, dim3(gx, nt), dim3(256), 0, s, descs, nt);
  else
    hipLaunchKernelGGL(k_theta<64>, dim3(gx, nt), dim3(256), 0, s, descs, nt);
}
void launch_winit(const UpdDesc* descs, int nt, int gx, hipStream_t s) {
  hipLaunchKernelGGL(k_winit, dim3(gx, nt), dim3(256), 0, s, descs);
}
void launch_finish(const UpdDesc* descs, int nt, int sweeps_run, hipStream_t s) {
  hipLaunchKernelGGL(k_finish, dim3(nt), dim3(1024), 0, s, descs, nt, sweeps_run);
}
void launch_fmul(const UpdDesc* descs, int nt, int maxX, int maxRank, hipStream_t s) {
  if (nt <= 4)
    hipLaunchKernelGGL(k_fmul<2>, dim3((maxX + 63) / 64, (maxRank + 7) / 8, nt), dim3(256), 0,
                       s, descs);
  else
    hipLaunchKernelGGL(k_fmul<8>, dim3((maxX + 63) / 64, (maxRank + 31) / 32, nt), dim3(256), 0,
                       s, descs);
}
void launch_split(const UpdDesc* descs, int nt, hipStream_t s) {
  hipLaunchKernelGGL(k_split, dim3(kNPart, nt), dim3(256), 0, s, descs);
}
void launch_renorm(const UpdDesc* descs, int nt, hipStream_t s) {
  hipLaunchKernelGGL(k_renorm, dim3(kNPart, nt), dim3(256), 0, s, descs);
}
void launch_measure(const MeasDesc* descs, int nm, int G, double* parts, hipStream_t s) {
  hipLaunchKernelGGL(k_measure_part, dim3(G, nm), dim3(256), 0, s, descs, parts);
  hipLaunchKernelGGL(k_measure_project, dim3(G, nm), dim3(256), 0, s, descs,
                     (const double*)parts);
}
void launch_entropy(const EntDesc* descs, int count, double* out, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_entropy, dim3(count), dim3(256), 0, s, descs, out);
}
void launch_contract(const double2* acc, size_t P, const int* dims, const double2* g,
                     const double* lamL, size_t max_out, double2* out, hipStream_t s) {
  const size_t blocks = std::min<size_t>((max_out + 255) / 256, 65536);
  hipLaunchKernelGGL(k_contract, dim3((unsigned)std::max<size_t>(blocks, 1)), dim3(256), 0, s,
                     acc, P, dims, g, lamL, out);
}
void launch_rho_part(const double2* t, size_t L, size_t D, size_t R, const double* lamR,
                     double2* rho, hipStream_t s) {
  const unsigned g = (unsigned)((D + 15) / 16);
  hipLaunchKernelGGL(k_rho_part, dim3(g, g), dim3(256), 0, s, t, D, R, L * R, lamR, rho);
}
void launch_fv_init(const FvArgs& a, hipStream_t s) {
  const size_t work = std::max((size_t)a.C * a.C, a.orient == 1 ? (size_t)a.m * a.ncols : 1);
  const int gx = (int)std::min<size_t>(std::max<size_t>((work + 255) / 256, 1), 2048);
  hipLaunchKernelGGL(k_fv_init, dim3(gx), dim3(256), 0, s, a);
}
void launch_fv_split(const FvArgs& a, hipStream_t s) {
  const size_t work = std::max((size_t)a.m * a.rank, (size_t)a.rank * a.ncols);
  const int gx = (int)std::min<size_t>(std::max<size_t>((work + 255) / 256, 1), 2048);
  hipLaunchKernelGGL(k_fv_split, dim3(gx), dim3(256), 0, s, a);
}
void launch_fv_last(const double2* q, const double* lam, int udim, double2* g, hipStream_t s) {
  hipLaunchKernelGGL(k_fv_last, dim3((2 * udim + 255) / 256), dim3(256), 0, s, q, lam, udim, g);
}
void launch_u1(double2* g, const int* dims, const double* op8, int gx, hipStream_t s) {
  Gate2 op;
  for (int i = 0; i < 4; ++i) op.m[i] = make_double2(op8[2 * i], op8[2 * i + 1]);
  hipLaunchKernelGGL(k_u1, dim3(gx), dim3(256), 0, s, g, dims, op);
}

// ---------------------------------------------------------------- haar
// One thread per unitary u: counter (block, ids[3u], ids[3u+1], ids[3u+2]),
// key (seed, 0), blocks 0..7 give 32 uniforms, Box-Muller pairs give the
// normals Re z_e, Im z_e of entry e = 4 i + j of Z (row-major), and
// Gram-Schmidt on Z's columns (each projection applied twice) gives the Q of
// Z = Q R with R's diagonal real and positive.
__global__ __launch_bounds__(64) void k_haar4(uint64_t seed, const uint64_t* __restrict__ ids,
                                              int count, double2* __restrict__ out) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= count) return;
  const uint64_t c1 = ids[3 * (size_t)u], c2 = ids[3 * (size_t)u + 1],
                 c3 = ids[3 * (size_t)u + 2];
  double2 q[4][4];  // q[column][row]
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const Philox4x64 w = philox4x64_10((uint64_t)b, c1, c2, c3, seed, 0);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const double rad = sqrt(-2.0 * log(1.0 - u01(w.v[2 * h])));
      double sn, cs;
      sincospi(2.0 * u01(w.v[2 * h + 1]), &sn, &cs);
      const int e = 2 * b + h;  // entry (e / 4, e % 4)
      q[e % 4][e / 4] = make_double2(rad * cs, rad * sn);
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
      for (int i = 0; i < j; ++i) {
        double2 c = make_double2(0.0, 0.0);
#pragma unroll
        for (int r = 0; r < 4; ++r) c = cadd(c, cmulc(q[i][r], q[j][r]));
#pragma unroll
        for (int r = 0; r < 4; ++r) q[j][r] = cadd(q[j][r], cmul(make_double2(-c.x, -c.y), q[i][r]));
      }
    }
    double nrm2 = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) nrm2 += cabs2(q[j][r]);
    const double inv = 1.0 / sqrt(nrm2);
#pragma unroll
    for (int r = 0; r < 4; ++r) q[j][r] = cscale(q[j][r], inv);
  }
  double2* o = out + 16 * (size_t)u;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) o[4 * i + j] = q[j][i];
}

void launch_haar4(uint64_t seed, const uint64_t* ids, int count, double2* out, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_haar4, dim3((count + 63) / 64), dim3(64), 0, s, seed, ids, count, out);
}

}  // namespace tnmps
