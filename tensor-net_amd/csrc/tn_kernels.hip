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
//   k_winit        W = I for the Jacobi SVD
//   k_jacobi_step  one round-robin step of a block one-sided Jacobi SVD
//                  (stands in for zgesvd, mps.rs:299)
//   k_finish       singular values, sort, rank rule, renormalisation
//                  (process_svd, mps.rs:236-293)
//   k_fmul         factored updates: kept singular vectors times the
//                  isometric factor
//   k_split        new Gamma[k], Gamma[k+1] with lambda divisions and partial
//                  norms (mps.rs:852-876, 1357-1380, local_norm 1147-1202)
//   k_renorm       local_renormalize of both sites (mps.rs:1206-1209)
//   k_measure      Born probabilities, sampling or forced outcome, projection
//                  (mps.rs:1027-1134, 1427-1465, 2229-2242)
//   k_entropy      von Neumann / Renyi entropy of one bond (mps.rs:1251-1282)
//   k_u1           single-site unitary (mps.rs:1298-1314)
#include "tn_device.h"
#include "tn_kernels.h"

#include <hip/hip_cooperative_groups.h>

#include <cfloat>

namespace cg = cooperative_groups;

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
__device__ double block_sum(double v, double* buf) {
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

__device__ double block_max(double v, double* buf) {
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

// ---------------------------------------------------------------- theta
template <typename C2, typename R>
__device__ __forceinline__ C2 tc(double2 a) {
  C2 r;
  r.x = (R)a.x;
  r.y = (R)a.y;
  return r;
}
template <typename C2, typename R>
__device__ __forceinline__ C2 tmad(C2 acc, C2 a, C2 b) {
  acc.x += a.x * b.x - a.y * b.y;
  acc.y += a.x * b.y + a.y * b.x;
  return acc;
}

// All four (s,t) values of theta'[v,(s,t),w] in arithmetic C2/R (FP64 or
// FP32 with FP32 accumulation); the result is widened to double.
template <typename C2, typename R>
__device__ __forceinline__ void theta_point(const UpdDesc& d, int v, int w, int chiM, int chiR,
                                            double2 out[4]) {
  C2 acc[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) acc[i].x = acc[i].y = (R)0;
  const double2* __restrict__ a0 = d.gk + (size_t)(v * 2 + 0) * chiM;
  const double2* __restrict__ a1 = d.gk + (size_t)(v * 2 + 1) * chiM;
  for (int u = 0; u < chiM; ++u) {
    const R lm = (R)d.lamM[u];
    C2 b0 = tc<C2, R>(d.gk1[(size_t)(u * 2 + 0) * chiR + w]);
    C2 b1 = tc<C2, R>(d.gk1[(size_t)(u * 2 + 1) * chiR + w]);
    b0.x *= lm; b0.y *= lm;
    b1.x *= lm; b1.y *= lm;
    const C2 x0 = tc<C2, R>(a0[u]), x1 = tc<C2, R>(a1[u]);
    acc[0] = tmad<C2, R>(acc[0], x0, b0);  // s'=0 t'=0
    acc[1] = tmad<C2, R>(acc[1], x0, b1);  // s'=0 t'=1
    acc[2] = tmad<C2, R>(acc[2], x1, b0);  // s'=1 t'=0
    acc[3] = tmad<C2, R>(acc[3], x1, b1);  // s'=1 t'=1
  }
  if (d.lamL) {
    const R ll = (R)d.lamL[v];
#pragma unroll
    for (int i = 0; i < 4; ++i) { acc[i].x *= ll; acc[i].y *= ll; }
  }
  if (d.lamR) {
    const R lr = (R)d.lamR[w];
#pragma unroll
    for (int i = 0; i < 4; ++i) { acc[i].x *= lr; acc[i].y *= lr; }
  }
  if (d.has_gate) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      C2 o;
      o.x = o.y = (R)0;
#pragma unroll
      for (int j = 0; j < 4; ++j) o = tmad<C2, R>(o, tc<C2, R>(d.gate[i * 4 + j]), acc[j]);
      out[i] = make_double2((double)o.x, (double)o.y);
    }
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) out[i] = make_double2((double)acc[i].x, (double)acc[i].y);
  }
}

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

// theta'[v,(s,t),w] = sum_{s't'} U[(s,t),(s',t')] lamL_v lamR_w
//                     sum_u Gk[v,s',u] lamM_u Gk1[u,t',w]
// written into X as theta (orient 0, m >= n) or theta^H (orient 1).
// Tile: 4 values of v x 64 values of w per 256-thread block.  In the
// factored modes X is filled elementwise instead (no contraction):
//   mode 1: X[(t,w), u] = conj(lamM_u Gk1[u,t,w] lamR_w)
//   mode 2: X[(v,s), u] = lamL_v Gk[v,s,u] lamM_u
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
    const size_t tot = (size_t)ts.R * chiM;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += stride) {
      if (ts.mode == 1) {
        const int u = (int)(e / ts.n), r = (int)(e % ts.n);
        double f = d.lamM[u];
        if (d.lamR) f *= d.lamR[r % chiR];
        d.X[(size_t)u * d.ldx + r] = cconj(cscale(d.gk1[e], f));
      } else {
        const int r = (int)(e / chiM), u = (int)(e % chiM);
        double f = d.lamM[u];
        if (d.lamL) f *= d.lamL[r >> 1];
        d.X[(size_t)u * d.ldx + r] = cscale(d.gk[e], f);
      }
    }
    return;
  }
  const int nvt = (chiL + 3) / 4, nwt = (chiR + 63) / 64;
  const int tv = threadIdx.x >> 6, tw = threadIdx.x & 63;
  for (int tile = blockIdx.x; tile < nvt * nwt; tile += gridDim.x) {
    const int v = (tile / nwt) * 4 + tv;
    const int w = (tile % nwt) * 64 + tw;
    if (v >= chiL || w >= chiR) continue;
    double2 out[4];
    if (d.prec == 1)
      theta_point<float2, float>(d, v, w, chiM, chiR, out);
    else
      theta_point<double2, double>(d, v, w, chiM, chiR, out);
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

// ---------------------------------------------------------------- W = I
__global__ __launch_bounds__(256) void k_winit(const UpdDesc* __restrict__ descs) {
  const UpdDesc& d = descs[blockIdx.y];
  const int C = task_shape(d).C;
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < d.pcnt_cap; i += blockDim.x) d.pcnt[i] = 0u;
  const size_t tot = (size_t)C * C;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
       e += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % C), j = (int)(e / C);
    d.W[(size_t)j * d.ldw + i] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
}

// ---------------------------------------------------------- Jacobi step
// Block one-sided Jacobi (Hestenes) over column blocks of width BW (8 or 16).
// A sweep visits every pair of blocks once, in the round-robin ("circle")
// ordering of the nbe blocks (nbe = number of blocks rounded up to even; a
// ghost block holds no columns); one step holds nbe/2 disjoint pairs.  For its
// pair (a, b) a workgroup forms the panel Gram matrix G = P^H P (PW x PW,
// PW = 2 BW), diagonalises it with cyclic two-sided Jacobi rotations
// (G <- Q^H G Q, Wl <- Wl Q), and applies Wl to the panel of X and of W.  The
// largest relative off-diagonal |G_ij| / sqrt(G_ii G_jj) seen in a sweep goes
// to offs[sweep]; a task whose previous sweep stayed at or below conv_tol has
// converged.  BW = 16 halves the steps per sweep; it serves lone large SVDs,
// whose time is set by the chain of steps rather than by the flops.
__device__ __forceinline__ int rr_player(int i, int step, int nbe) {
  return 1 + (i - 1 + step) % (nbe - 1);
}

// Columns whose squared norm is below kTinyNorm2 (norm 1e-17, under
// f64::EPSILON times a unit-norm theta, so below every truncation threshold of
// the rank rule) take part in no rotation and in no convergence measure.
constexpr double kTinyNorm2 = 1e-34;

template <int BW>
struct PairSmem {
  static constexpr int PW = 2 * BW, LD = PW + 1;
  double2 Ps[64][LD];           // 64-row chunk of the panel
  double2 G[PW][LD];            // panel Gram matrix
  double2 Wl[PW][LD];           // its accumulated rotation
  double2 scratch[4 * PW * LD]; // per-wave partial Gram, then Gb[2] + Wb[2]
  int rr_partner[PW - 1][PW];   // round-robin partner of each player
  int rr_pair[PW - 1][PW];      // index of the player's pair in the round
  double rcs[PW / 2];           // rotation of each pair of the round
  double2 rsn[PW / 2];
  int col[PW];
  double red[4];
  __device__ double2& part(int w, int i, int j) { return scratch[(w * PW + i) * LD + j]; }
  __device__ double2& Gb(int b, int i, int j) { return scratch[(b * PW + i) * LD + j]; }
  __device__ double2& Wb(int b, int i, int j) { return scratch[((2 + b) * PW + i) * LD + j]; }
};

// Columns of block pair p at round-robin step `step` into sm.col (-1 for
// columns beyond C).  False (for the whole block) when the pair does not
// exist for this task's C.
template <int BW>
__device__ bool pair_cols(int C, int step, int p, PairSmem<BW>& sm) {
  constexpr int PW = 2 * BW;
  const int nb = (C + BW - 1) / BW;
  const int nbe = nb + (nb & 1);
  if (step >= nbe - 1 || p >= nbe / 2) return false;
  const int ba = (p == 0) ? 0 : rr_player(p, step, nbe);
  const int bb = rr_player(nbe - 1 - p, step, nbe);
  const int tid = threadIdx.x;
  if (tid < PW) {
    const int c = (tid < BW) ? ba * BW + tid : bb * BW + (tid - BW);
    sm.col[tid] = (c < C) ? c : -1;
  }
  __syncthreads();
  return true;
}

// Stage rows [r0, min(r0 + 64, r_end)) of the panel of M into sm.Ps (zeros
// past r_end or for absent columns).
template <int BW>
__device__ __forceinline__ void stage_rows(const double2* __restrict__ M, size_t ld, int r0,
                                           int r_end, PairSmem<BW>& sm) {
  constexpr int PW = 2 * BW;
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < PW / 4; ++q) {
    const int e = tid + 256 * q;
    const int rr = e & 63, cc = e >> 6;
    const int c = sm.col[cc];
    sm.Ps[rr][cc] = (c >= 0 && r0 + rr < r_end) ? M[(size_t)c * ld + r0 + rr]
                                                : make_double2(0.0, 0.0);
  }
}

// sm.G = P^H P over rows [r_begin, r_end) of the panel of X.  Thread = (row
// slice rs, 4x4 tile of G); a row slice takes every RS-th row of each
// 64-row chunk.  Slices are reduced by shuffles (inside a wave) and through
// LDS (across the four waves).
template <int BW>
__device__ void panel_gram(const UpdDesc& d, int r_begin, int r_end, PairSmem<BW>& sm) {
  constexpr int PW = 2 * BW, TPR = PW / 4, NTILE = TPR * TPR, RS = 256 / NTILE;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int tile = tid % NTILE, rs = tid / NTILE;
  const int ti = tile / TPR, tj = tile % TPR;
  double2 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = make_double2(0.0, 0.0);
  for (int r0 = r_begin; r0 < r_end; r0 += 64) {
    stage_rows<BW>(d.X, (size_t)d.ldx, r0, r_end, sm);
    __syncthreads();
#pragma unroll 2
    for (int k = 0; k < 64 / RS; ++k) {
      const int rr = rs + RS * k;
      double2 xi[4], xj[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) xi[a] = sm.Ps[rr][ti * 4 + a];
#pragma unroll
      for (int b = 0; b < 4; ++b) xj[b] = sm.Ps[rr][tj * 4 + b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = cadd(acc[a][b], cmulc(xi[a], xj[b]));
    }
    __syncthreads();
  }
  // lanes tile, tile + NTILE, ... of one wave hold the same tile
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      double2 v = acc[a][b];
#pragma unroll
      for (int off = NTILE; off < 64; off <<= 1) {
        v.x += __shfl_xor(v.x, off, 64);
        v.y += __shfl_xor(v.y, off, 64);
      }
      acc[a][b] = v;
    }
  if (lane < NTILE) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) sm.part(wid, ti * 4 + a, tj * 4 + b) = acc[a][b];
  }
  __syncthreads();
  for (int e = tid; e < PW * PW; e += 256) {
    const int gi = e / PW, gj = e % PW;
    double2 g = sm.part(0, gi, gj);
    g = cadd(g, sm.part(1, gi, gj));
    g = cadd(g, sm.part(2, gi, gj));
    g = cadd(g, sm.part(3, gi, gj));
    sm.G[gi][gj] = g;
  }
  __syncthreads();
}

// 1/sqrt(x) and 1/x from the hardware approximations refined by two Newton
// steps each (quadratic convergence to double precision): cheaper than the
// IEEE-exact division and square-root sequences on the rotation critical path.
__device__ __forceinline__ double rsq_nr(double x) {
  double r = __builtin_amdgcn_rsq(x);
  r = r * (1.5 - 0.5 * x * r * r);
  r = r * (1.5 - 0.5 * x * r * r);
  return r;
}
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = r * (2.0 - x * r);
  r = r * (2.0 - x * r);
  return r;
}

// Jacobi rotation zeroing the off-diagonal of [[a, c], [conj(c), b]]:
// Q = [[cs, sn], [-conj(sn), cs]] with sn = s e^{i arg c} (Golub-Van Loan
// symmetric Schur 2x2 with the phase of c factored out).  Returns false
// (identity) when the pair is already orthogonal to tol or a column is
// negligible.
__device__ __forceinline__ bool rot2(double a, double b, double2 c, double tol, double& cs,
                                     double2& sn) {
  const double c2 = cabs2(c);
  if (!(a > kTinyNorm2 && b > kTinyNorm2 && c2 > tol * tol * a * b)) {
    cs = 1.0;
    sn = make_double2(0.0, 0.0);
    return false;
  }
  const double rc = rsq_nr(c2);  // 1/|c|
  const double tau = (b - a) * 0.5 * rc;
  const double x = 1.0 + tau * tau;
  const double t = (tau >= 0.0 ? 1.0 : -1.0) * rcp_nr(fabs(tau) + x * rsq_nr(x));
  cs = rsq_nr(1.0 + t * t);
  const double f = t * cs * rc;
  sn = make_double2(f * c.x, f * c.y);
  return true;
}

// From the panel Gram matrix in sm.G: record the largest relative
// off-diagonal (when `record`), and unless the panel is already orthogonal
// diagonalise G with cyclic two-sided Jacobi rotations (round-robin order,
// PW/2 disjoint rotations per round; in each pair the lower index plays the
// first role), accumulating the rotation in sm.Wl.  Returns false (for the
// whole block) when there is nothing to apply.  A function of sm.G alone:
// blocks holding the same G compute bitwise the same Wl.
template <int BW>
__device__ bool panel_solve(const UpdDesc& d, int nt, int sweep, bool record, double tol_rot,
                            PairSmem<BW>& sm) {
  constexpr int PW = 2 * BW;
  const int tid = threadIdx.x;
  double ratio = 0.0;
  for (int e = tid; e < PW * PW; e += 256) {
    const int gi = e / PW, gj = e % PW;
    if (gi < gj) {
      const double a = sm.G[gi][gi].x, b = sm.G[gj][gj].x;
      if (a > kTinyNorm2 && b > kTinyNorm2)
        ratio = fmax(ratio, sqrt(cabs2(sm.G[gi][gj])) / sqrt(a * b));
    }
  }
  const double off = block_max(ratio, sm.red);
  if (record && tid == 0 && off > 0.0)
    atomicMax((unsigned long long*)&d.offs[sweep * nt],
              (unsigned long long)__double_as_longlong(off));
  if (off <= tol_rot) return false;

  for (int e = tid; e < (PW - 1) * PW; e += 256) {
    const int r = e / PW, x = e % PW;
    const int pos = (x == 0) ? 0 : 1 + ((x - 1 - r) % (PW - 1) + (PW - 1)) % (PW - 1);
    const int other = PW - 1 - pos;
    sm.rr_partner[r][x] = (other == 0) ? 0 : rr_player(other, r, PW);
    sm.rr_pair[r][x] = pos < PW / 2 ? pos : other;
  }
  for (int e = tid; e < PW * PW; e += 256) {
    const int gi = e / PW, gj = e % PW;
    sm.Gb(0, gi, gj) = sm.G[gi][gj];
    sm.Wb(0, gi, gj) = make_double2(gi == gj ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  int cur = 0;
  for (int isw = 0; isw < d.inner_sweeps; ++isw) {
    int rotated = 0;
    for (int round = 0; round < PW - 1; ++round) {
      if (tid < PW / 2) {
        const int p1 = (tid == 0) ? 0 : rr_player(tid, round, PW);
        const int p2 = rr_player(PW - 1 - tid, round, PW);
        const int i1 = min(p1, p2), i2 = max(p1, p2);
        double cs;
        double2 sn;
        rotated |= rot2(sm.Gb(cur, i1, i1).x, sm.Gb(cur, i2, i2).x, sm.Gb(cur, i1, i2), tol_rot,
                        cs, sn);
        sm.rcs[tid] = cs;
        sm.rsn[tid] = sn;
      }
      __syncthreads();
      for (int e = tid; e < PW * PW; e += 256) {
        const int gi = e / PW, gj = e % PW;
        const int qi = sm.rr_partner[round][gi], qj = sm.rr_partner[round][gj];
        const int pi = sm.rr_pair[round][gi], pj = sm.rr_pair[round][gj];
        // column j of Q: Q[j][j] = cs; off entry Q[qj][j] = -conj(sn) when j
        // is the lower index of its pair, sn when it is the upper one
        const double2 snj = sm.rsn[pj], sni = sm.rsn[pi];
        const double2 Qjj = make_double2(sm.rcs[pj], 0.0);
        const double2 Qoj = (gj < qj) ? make_double2(-snj.x, snj.y) : snj;
        const double2 Qii = make_double2(sm.rcs[pi], 0.0);
        const double2 Qoi = (gi < qi) ? make_double2(-sni.x, sni.y) : sni;
        const double2 g_ij = sm.Gb(cur, gi, gj), g_ioj = sm.Gb(cur, gi, qj);
        const double2 g_oij = sm.Gb(cur, qi, gj), g_oioj = sm.Gb(cur, qi, qj);
        const double2 t0 = cadd(cmul(g_ij, Qjj), cmul(g_ioj, Qoj));
        const double2 t1 = cadd(cmul(g_oij, Qjj), cmul(g_oioj, Qoj));
        sm.Gb(cur ^ 1, gi, gj) = cadd(cmulc(Qii, t0), cmulc(Qoi, t1));
        sm.Wb(cur ^ 1, gi, gj) =
            cadd(cmul(sm.Wb(cur, gi, gj), Qjj), cmul(sm.Wb(cur, gi, qj), Qoj));
      }
      cur ^= 1;
      __syncthreads();
    }
    if (!__syncthreads_or(rotated)) break;
  }
  for (int e = tid; e < PW * PW; e += 256) {
    const int gi = e / PW, gj = e % PW;
    sm.Wl[gi][gj] = sm.Wb(cur, gi, gj);
  }
  __syncthreads();
  return true;
}

// Rows [a, b) of the panel of M (leading dimension ld) <- panel * Wl.
// Thread = (row lane 0..63, output group jg 0..3 of PW/4 columns); rows are
// staged through LDS 64 at a time.
template <int BW>
__device__ void panel_apply(double2* M, size_t ld, int a, int b, PairSmem<BW>& sm) {
  constexpr int PW = 2 * BW, NJ = PW / 4;
  const int tid = threadIdx.x;
  const int rl = tid & 63, jg = tid >> 6;
  for (int r0 = a; r0 < b; r0 += 64) {
    stage_rows<BW>(M, ld, r0, b, sm);
    __syncthreads();
    const int r = r0 + rl;
    double2 o[NJ];
#pragma unroll
    for (int k = 0; k < NJ; ++k) o[k] = make_double2(0.0, 0.0);
#pragma unroll 4
    for (int c = 0; c < PW; ++c) {
      const double2 x = sm.Ps[rl][c];
#pragma unroll
      for (int k = 0; k < NJ; ++k) o[k] = cadd(o[k], cmul(x, sm.Wl[c][jg * NJ + k]));
    }
    if (r < b) {
#pragma unroll
      for (int k = 0; k < NJ; ++k) {
        const int c = sm.col[jg * NJ + k];
        if (c >= 0) M[(size_t)c * ld + r] = o[k];
      }
    }
    __syncthreads();
  }
}

// Process block pair p of step `step` of sweep `sweep` for task d with one
// workgroup.  Returns normally in every case (no barrier is skipped by a
// subset of the block).
template <int BW>
__device__ void jacobi_pair(const UpdDesc& d, int nt, int sweep, int step, int p,
                            double tol_rot, PairSmem<BW>& sm) {
  const int R = d.info[kInfoR], C = d.info[kInfoC];
  if (!pair_cols<BW>(C, step, p, sm)) return;
  panel_gram<BW>(d, 0, R, sm);
  if (!panel_solve<BW>(d, nt, sweep, true, tol_rot, sm)) return;
  panel_apply<BW>(d.X, (size_t)d.ldx, 0, R, sm);
  panel_apply<BW>(d.W, (size_t)d.ldw, 0, C, sm);
}

// Row range of slice q of S over n rows (slices are multiples of 64 rows).
__device__ __forceinline__ void split_rows(int n, int S, int q, int& a, int& b) {
  const int per = (((n + S - 1) / S) + 63) / 64 * 64;
  a = min(n, q * per);
  b = min(n, a + per);
}

// Row-split step, part 1: workgroup (p, q) writes the Gram matrix of rows
// slice q of pair p's panel to gpart[p][q].
template <int BW>
__global__ __launch_bounds__(256, BW == 8 ? 2 : 1) void k_jacobi_gram(const UpdDesc* __restrict__ descs, int nt,
                                                     int sweep, int step, int S,
                                                     double conv_tol) {
  constexpr int PW = 2 * BW;
  __shared__ PairSmem<BW> sm;
  const UpdDesc& d = descs[blockIdx.y];
  if (sweep > 0 && d.offs[(sweep - 1) * nt] <= conv_tol) return;
  const int R = d.info[kInfoR], C = d.info[kInfoC];
  const int p = blockIdx.x / S, q = blockIdx.x - (blockIdx.x / S) * S;
  if (!pair_cols<BW>(C, step, p, sm)) return;
  int a, b;
  split_rows(R, S, q, a, b);
  panel_gram<BW>(d, a, b, sm);
  for (int e = threadIdx.x; e < PW * PW; e += 256)
    d.gpart[((size_t)p * S + q) * (PW * PW) + e] = sm.G[e / PW][e % PW];
}

// Row-split step, part 2: every workgroup of pair p sums the S partial Gram
// matrices in the same order, so all of them solve the same G into the same
// Wl, and applies it to its own slice of the rows of X and of W.
template <int BW>
__global__ __launch_bounds__(256, BW == 8 ? 2 : 1) void k_jacobi_apply(const UpdDesc* __restrict__ descs, int nt,
                                                      int sweep, int step, int S,
                                                      double tol_rot, double conv_tol) {
  constexpr int PW = 2 * BW;
  __shared__ PairSmem<BW> sm;
  const UpdDesc& d = descs[blockIdx.y];
  if (sweep > 0 && d.offs[(sweep - 1) * nt] <= conv_tol) return;
  const int R = d.info[kInfoR], C = d.info[kInfoC];
  const int p = blockIdx.x / S, q = blockIdx.x - (blockIdx.x / S) * S;
  if (!pair_cols<BW>(C, step, p, sm)) return;
  for (int e = threadIdx.x; e < PW * PW; e += 256) {
    double2 g = make_double2(0.0, 0.0);
    for (int s = 0; s < S; ++s) g = cadd(g, d.gpart[((size_t)p * S + s) * (PW * PW) + e]);
    sm.G[e / PW][e % PW] = g;
  }
  __syncthreads();
  if (!panel_solve<BW>(d, nt, sweep, q == 0, tol_rot, sm)) return;
  int a, b;
  split_rows(R, S, q, a, b);
  panel_apply<BW>(d.X, (size_t)d.ldx, a, b, sm);
  split_rows(C, S, q, a, b);
  panel_apply<BW>(d.W, (size_t)d.ldw, a, b, sm);
}

// Fused row-split step (cooperative launch: every workgroup of the grid is
// resident): workgroup (p, q) writes the Gram matrix of its row slice, meets
// the other S-1 workgroups of pair p at a pair barrier (an arrival counter in
// device memory; epoch = the task's step count so far), then sums the S
// partials in the same order as its partners, solves and applies to its
// rows.  The spin is bounded: past kSpinLimit polls the workgroup raises
// abort_flag and every later wait returns at once, so a failed barrier ends
// the kernel with wrong results instead of hanging it.
constexpr unsigned kSpinLimit = 1u << 22;

__device__ __forceinline__ void pair_barrier(unsigned* cnt, unsigned target, int* abort_flag) {
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(cnt, 1u);
    unsigned spins = 0;
    while (__hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
      __builtin_amdgcn_s_sleep(1);
      if (++spins > kSpinLimit) {
        atomicExch(abort_flag, 1);
        break;
      }
    }
  }
  __syncthreads();
  __threadfence();
}

template <int BW>
__global__ __launch_bounds__(256, BW == 8 ? 2 : 1) void k_jacobi_fused(
    const UpdDesc* __restrict__ descs, int nt, int sweep, int step, int S, double tol_rot,
    double conv_tol) {
  constexpr int PW = 2 * BW;
  __shared__ PairSmem<BW> sm;
  const UpdDesc& d = descs[blockIdx.y];
  if (sweep > 0 && d.offs[(sweep - 1) * nt] <= conv_tol) return;
  const int R = d.info[kInfoR], C = d.info[kInfoC];
  const int p = blockIdx.x / S, q = blockIdx.x - (blockIdx.x / S) * S;
  if (!pair_cols<BW>(C, step, p, sm)) return;
  const int nb = (C + BW - 1) / BW;
  const int nbe = nb + (nb & 1);
  const unsigned epoch = (unsigned)sweep * (unsigned)(nbe - 1) + (unsigned)step;
  int a, b;
  split_rows(R, S, q, a, b);
  panel_gram<BW>(d, a, b, sm);
  for (int e = threadIdx.x; e < PW * PW; e += 256)
    d.gpart[((size_t)p * S + q) * (PW * PW) + e] = sm.G[e / PW][e % PW];
  pair_barrier(&d.pcnt[p], (unsigned)S * (epoch + 1u), d.abort_flag);
  for (int e = threadIdx.x; e < PW * PW; e += 256) {
    double2 g = make_double2(0.0, 0.0);
    for (int s = 0; s < S; ++s) g = cadd(g, d.gpart[((size_t)p * S + s) * (PW * PW) + e]);
    sm.G[e / PW][e % PW] = g;
  }
  __syncthreads();
  if (!panel_solve<BW>(d, nt, sweep, q == 0, tol_rot, sm)) return;
  panel_apply<BW>(d.X, (size_t)d.ldx, a, b, sm);
  split_rows(C, S, q, a, b);
  panel_apply<BW>(d.W, (size_t)d.ldw, a, b, sm);
}

// One launch per step, one workgroup per block pair.
template <int BW>
__global__ __launch_bounds__(256, BW == 8 ? 2 : 1) void k_jacobi_step(const UpdDesc* __restrict__ descs, int nt,
                                                     int sweep, int step, double tol_rot,
                                                     double conv_tol) {
  __shared__ PairSmem<BW> sm;
  const UpdDesc& d = descs[blockIdx.y];
  if (sweep > 0 && d.offs[(sweep - 1) * nt] <= conv_tol) return;
  jacobi_pair<BW>(d, nt, sweep, step, blockIdx.x, tol_rot, sm);
}

// Persistent cooperative kernel: all sweeps and steps of a batch of SVDs in
// one launch.  Work item = (task, pair); a grid barrier separates steps.
// Every block runs the same sequence of grid barriers: the loops have
// grid-uniform bounds and the convergence decision after each sweep is
// computed by every block from the same (final) per-sweep maxima.
__global__ __launch_bounds__(256) void k_jacobi_coop(const UpdDesc* __restrict__ descs, int nt,
                                                     int npairs, int steps, int max_sweeps,
                                                     double tol_rot, double conv_tol,
                                                     int* sweeps_out) {
  __shared__ PairSmem<8> sm;
  __shared__ int s_done;
  cg::grid_group grid = cg::this_grid();
  const int items = nt * npairs;
  for (int sw = 0; sw < max_sweeps; ++sw) {
    for (int st = 0; st < steps; ++st) {
      for (int it = blockIdx.x; it < items; it += gridDim.x) {
        const int t = it / npairs, p = it - t * npairs;
        const UpdDesc& d = descs[t];
        if (sw > 0 && d.offs[(sw - 1) * nt] <= conv_tol) continue;
        jacobi_pair<8>(d, nt, sw, st, p, tol_rot, sm);
      }
      grid.sync();
    }
    if (threadIdx.x == 0) {
      int done = 1;
      for (int t = 0; t < nt; ++t)
        if (!(descs[t].offs[sw * nt] <= conv_tol)) {
          done = 0;
          break;
        }
      s_done = done;
      if (blockIdx.x == 0) *sweeps_out = sw + 1;
    }
    __syncthreads();
    const int done = s_done;
    __syncthreads();
    if (done) break;
  }
}

// ------------------------------------------------------------- finish
constexpr double kF32NoiseFloor = 9.5367431640625e-07;  // 2^-20

// Column norms of X are the singular values (theta = U S V^H).  Sort them
// descending, apply process_svd (mps.rs:236-293): normalise, pick the rank,
// renormalise, and record the right-factor scale fac_j = s_j * norm2.
__global__ __launch_bounds__(1024) void k_finish(const UpdDesc* __restrict__ descs, int nt,
                                                 const int* sweeps_dev, int sweeps_run) {
  const UpdDesc& d = descs[blockIdx.x];
  const int R = d.info[kInfoR], C = d.info[kInfoC];
  __shared__ double key[kMaxCols];
  __shared__ int idx[kMaxCols];
  __shared__ double red[16];
  __shared__ int first_bad;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;

  if (tid == 0) d.info[kInfoSweeps] = sweeps_dev ? *sweeps_dev : sweeps_run;

  for (int j = wid; j < C; j += 16) {
    double s = 0.0;
    for (int r = lane; r < R; r += 64) s += cabs2(d.X[(size_t)j * d.ldx + r]);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) d.sraw[j] = sqrt(s);
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
  // bitonic sort: descending key, ascending index on ties
  for (int k = 2; k <= P2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < P2; i += blockDim.x) {
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
  // they are not counted as kept values of the rank rule.
  if (d.prec == 1) thr = fmax(thr, kF32NoiseFloor);
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
  }
}

// ------------------------------------------------------ factor product
// Factored modes, after k_finish: the kept singular vectors times the
// isometric factor, for the kept columns j < rank.
//   mode 1: T[j][row] = sum_u Gk[row, u] W[u, perm_j]          (row < m)
//           = Gamma[k] of lamL Gamma[k] U_M (already divided by lamL)
//   mode 2: T[j][c]   = sum_u conj(W[u, perm_j]) Gk1[u, c]     (c < n)
//           = rows of V_A^H Gamma[k+1] (already divided by lamR)
// Tile: 64 values of the free index x 32 values of j per 256-thread block,
// u streamed through LDS 16 at a time.  Grid (x tiles, j tiles, task).
__global__ __launch_bounds__(256) void k_fmul(const UpdDesc* __restrict__ descs) {
  const UpdDesc& d = descs[blockIdx.z];
  const int mode = d.info[kInfoMode];
  if (mode == 0) return;
  const int K = d.info[kInfoChiM];
  const int rank = d.info[kInfoRank];
  const int nx = (mode == 1) ? d.info[kInfoM] : d.info[kInfoN];
  const int x0 = blockIdx.x * 64, j0 = blockIdx.y * 32;
  if (x0 >= nx || j0 >= rank) return;
  __shared__ double2 Ps[16][65];  // [u][x]
  __shared__ double2 Ws[16][33];  // [u][j]
  __shared__ int pj[32];
  const int tid = threadIdx.x;
  if (tid < 32) pj[tid] = (j0 + tid < rank) ? d.perm[j0 + tid] : -1;
  __syncthreads();
  const int tx = tid & 63, ty = tid >> 6;
  double2 acc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = make_double2(0.0, 0.0);
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
    for (int q = 0; q < 2; ++q) {
      const int e = tid + 256 * q;
      const int uu = e & 15, jj = e >> 4;
      const int u = u0 + uu, c = pj[jj];
      double2 w = make_double2(0.0, 0.0);
      if (c >= 0 && u < K) w = d.W[(size_t)c * d.ldw + u];
      Ws[uu][jj] = (mode == 2) ? cconj(w) : w;
    }
    __syncthreads();
#pragma unroll 4
    for (int uu = 0; uu < 16; ++uu) {
      const double2 pv = Ps[uu][tx];
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] = cadd(acc[k], cmul(pv, Ws[uu][ty * 8 + k]));
    }
    __syncthreads();
  }
  const int x = x0 + tx;
  if (x < nx) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int j = j0 + ty * 8 + k;
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
  const size_t nk = (size_t)chiL * 2 * rank, nk1 = (size_t)rank * 2 * chiR;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  double accA = 0.0, accB = 0.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nk; e += stride) {
    const int j = (int)(e % rank);
    const int row = (int)(e / rank);  // v*2 + s
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
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nk1; e += stride) {
    const int j = (int)(e / (2 * (size_t)chiR));
    const int cidx = (int)(e % (2 * (size_t)chiR));  // t*chiR + w
    const int w = cidx % chiR;
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
  if (threadIdx.x < 2) {
    double s = 0.0;
    for (int b = 0; b < (int)gridDim.x; ++b) s += d.part[2 * b + threadIdx.x];
    nrm[threadIdx.x] = sqrt(s);
  }
  __syncthreads();
  const double nA = nrm[0], nB = nrm[1];
  const size_t nk = (size_t)chiL * 2 * rank, nk1 = (size_t)rank * 2 * chiR;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nk; e += stride)
    d.gk[e] = cdivr(d.gk[e], nA);
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nk1; e += stride)
    d.gk1[e] = cdivr(d.gk1[e], nB);
}

// ------------------------------------------------------------ measure
__global__ __launch_bounds__(1024) void k_measure(const MeasDesc* __restrict__ descs) {
  const MeasDesc& d = descs[blockIdx.x];
  const int chiL = d.dims[0], chiR = d.dims[1];
  __shared__ double red[16];
  __shared__ int s_out;
  __shared__ double s_prob;
  double p0 = 0.0, p1 = 0.0;
  const size_t tot = (size_t)chiL * chiR;
  for (size_t e = threadIdx.x; e < tot; e += blockDim.x) {
    const int v = (int)(e / chiR), u = (int)(e % chiR);
    double wgt = 1.0;
    if (d.lamL) wgt *= d.lamL[v] * d.lamL[v];
    if (d.lamR) wgt *= d.lamR[u] * d.lamR[u];
    p0 += cabs2(d.g[((size_t)v * 2 + 0) * chiR + u]) * wgt;
    p1 += cabs2(d.g[((size_t)v * 2 + 1) * chiR + u]) * wgt;
  }
  p0 = block_sum(p0, red);
  p1 = block_sum(p1, red);
  if (threadIdx.x == 0) {
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
    s_out = out;
    s_prob = prob;
    d.out_outcome[d.slot] = out;
    d.out_prob[d.slot] = prob;
    d.out_probs[2 * d.slot + 0] = p0;
    d.out_probs[2 * d.slot + 1] = p1;
    d.out_flag[d.slot] = flag;
  }
  __syncthreads();
  if (d.forced == -2) return;
  const int out = s_out;
  const double renorm = sqrt(s_prob);
  for (size_t e = threadIdx.x; e < tot; e += blockDim.x) {
    const int v = (int)(e / chiR), u = (int)(e % chiR);
    double2* keep = d.g + ((size_t)v * 2 + out) * chiR + u;
    double2* drop = d.g + ((size_t)v * 2 + (1 - out)) * chiR + u;
    *keep = cdivr(*keep, renorm);
    *drop = make_double2(0.0, 0.0);
  }
}

// ------------------------------------------------------------ entropy
// a == 1: -sum s^2 ln s^2; otherwise ln(sum (s^2)^a) / (1 - a).
__global__ __launch_bounds__(256) void k_entropy(const double* lam, const int* dimp, double a,
                                                 double* out) {
  __shared__ double red[4];
  const int len = *dimp;
  double acc = 0.0;
  for (int i = threadIdx.x; i < len; i += blockDim.x) {
    const double s2 = lam[i] * lam[i];
    if (a == 1.0) acc += -s2 * log(s2);
    else acc += pow(s2, a);
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) *out = (a == 1.0) ? s : log(s) * (1.0 / (1.0 - a));
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
void launch_theta(const UpdDesc* descs, int nt, int gx, hipStream_t s) {
  hipLaunchKernelGGL(k_theta, dim3(gx, nt), dim3(256), 0, s, descs, nt);
}
void launch_winit(const UpdDesc* descs, int nt, int gx, hipStream_t s) {
  hipLaunchKernelGGL(k_winit, dim3(gx, nt), dim3(256), 0, s, descs);
}
void launch_jacobi_step(const UpdDesc* descs, int nt, int npairs, int bw, int sweep, int step,
                        double tol_rot, double conv_tol, hipStream_t s) {
  if (bw == 16)
    hipLaunchKernelGGL(k_jacobi_step<16>, dim3(npairs, nt), dim3(256), 0, s, descs, nt, sweep,
                       step, tol_rot, conv_tol);
  else
    hipLaunchKernelGGL(k_jacobi_step<8>, dim3(npairs, nt), dim3(256), 0, s, descs, nt, sweep,
                       step, tol_rot, conv_tol);
}
void launch_jacobi_gram(const UpdDesc* descs, int nt, int npairs, int bw, int S, int sweep,
                        int step, double conv_tol, hipStream_t s) {
  if (bw == 16)
    hipLaunchKernelGGL(k_jacobi_gram<16>, dim3(npairs * S, nt), dim3(256), 0, s, descs, nt,
                       sweep, step, S, conv_tol);
  else
    hipLaunchKernelGGL(k_jacobi_gram<8>, dim3(npairs * S, nt), dim3(256), 0, s, descs, nt,
                       sweep, step, S, conv_tol);
}
void launch_jacobi_apply(const UpdDesc* descs, int nt, int npairs, int bw, int S, int sweep,
                         int step, double tol_rot, double conv_tol, hipStream_t s) {
  if (bw == 16)
    hipLaunchKernelGGL(k_jacobi_apply<16>, dim3(npairs * S, nt), dim3(256), 0, s, descs, nt,
                       sweep, step, S, tol_rot, conv_tol);
  else
    hipLaunchKernelGGL(k_jacobi_apply<8>, dim3(npairs * S, nt), dim3(256), 0, s, descs, nt,
                       sweep, step, S, tol_rot, conv_tol);
}
hipError_t launch_jacobi_fused(const UpdDesc* descs, int nt, int npairs, int S, int sweep,
                               int step, double tol_rot, double conv_tol, hipStream_t s) {
  void* args[] = {(void*)&descs, (void*)&nt, (void*)&sweep, (void*)&step, (void*)&S,
                  (void*)&tol_rot, (void*)&conv_tol};
  return hipLaunchCooperativeKernel((const void*)k_jacobi_fused<8>, dim3(npairs * S, nt),
                                    dim3(256), args, 0, s);
}
int jacobi_fused_max_blocks() {
  int dev = 0, per_cu = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
  if (!prop.cooperativeLaunch) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_jacobi_fused<8>, 256,
                                                   0) != hipSuccess)
    return 0;
  return per_cu * prop.multiProcessorCount;
}
hipError_t launch_jacobi_coop(const UpdDesc* descs, int nt, int npairs, int steps,
                              int max_sweeps, double tol_rot, double conv_tol, int* sweeps_out,
                              int grid, hipStream_t s) {
  void* args[] = {(void*)&descs, (void*)&nt, (void*)&npairs, (void*)&steps, (void*)&max_sweeps,
                  (void*)&tol_rot, (void*)&conv_tol, (void*)&sweeps_out};
  return hipLaunchCooperativeKernel((const void*)k_jacobi_coop, dim3(grid), dim3(256), args, 0,
                                    s);
}
int jacobi_coop_max_blocks() {
  int dev = 0, per_cu = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
  if (!prop.cooperativeLaunch) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_jacobi_coop, 256, 0) !=
      hipSuccess)
    return 0;
  return per_cu * prop.multiProcessorCount;
}
void launch_finish(const UpdDesc* descs, int nt, const int* sweeps_dev, int sweeps_run,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_finish, dim3(nt), dim3(1024), 0, s, descs, nt, sweeps_dev, sweeps_run);
}
void launch_fmul(const UpdDesc* descs, int nt, int maxX, int maxRank, hipStream_t s) {
  hipLaunchKernelGGL(k_fmul, dim3((maxX + 63) / 64, (maxRank + 31) / 32, nt), dim3(256), 0, s,
                     descs);
}
void launch_split(const UpdDesc* descs, int nt, hipStream_t s) {
  hipLaunchKernelGGL(k_split, dim3(kNPart, nt), dim3(256), 0, s, descs);
}
void launch_renorm(const UpdDesc* descs, int nt, hipStream_t s) {
  hipLaunchKernelGGL(k_renorm, dim3(kNPart, nt), dim3(256), 0, s, descs);
}
void launch_measure(const MeasDesc* descs, int nm, hipStream_t s) {
  hipLaunchKernelGGL(k_measure, dim3(nm), dim3(1024), 0, s, descs);
}
void launch_entropy(const double* lam, const int* dimp, double a, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_entropy, dim3(1), dim3(256), 0, s, lam, dimp, a, out);
}
void launch_u1(double2* g, const int* dims, const double* op8, int gx, hipStream_t s) {
  Gate2 op;
  for (int i = 0; i < 4; ++i) op.m[i] = make_double2(op8[2 * i], op8[2 * i + 1]);
  hipLaunchKernelGGL(k_u1, dim3(gx), dim3(256), 0, s, g, dims, op);
}

}  // namespace tnmps
