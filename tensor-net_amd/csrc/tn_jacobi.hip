// Block one-sided Jacobi SVD steps for gfx950 (stands in for LAPACK zgesvd,
// tensor-net/lib/mps.rs:299).
//
// Columns of X are grouped in blocks of BW (8 or 16).  A sweep visits every
// pair of blocks once, in the round-robin ("circle") ordering of nbe blocks
// (nbe = the widest task's block count rounded up to even; blocks past a
// task's C hold no columns); one step holds nbe/2 disjoint pairs.  For its
// pair (a, b) a workgroup forms the panel Gram matrix G = P^H P (PW x PW,
// PW = 2 BW), diagonalises it with cyclic two-sided Jacobi rotations
// (G <- Q^H G Q, Wl <- Wl Q), and applies Wl to the panel of X and of W.  The
// largest relative off-diagonal |G_ij| / sqrt(G_ii G_jj) seen in a sweep goes
// to offs[sweep]; a task whose previous sweep stayed at or below conv_tol has
// converged and its workgroups return at once.
//
// Launch shapes:
//   k_jacobi_step             one workgroup per block pair (many pairs)
//   k_jacobi_gram + _apply    S workgroups per pair, each on a slice of the
//                             rows: partial Gram matrices, then every one of
//                             the S sums them in the same order, solves the
//                             same G into bitwise the same Wl and applies it
//                             to its own slice (few pairs: a lone SVD)
//
// The kernels take JacobiArgs by value and compute every address from it
// and the task index, so their first global loads (the panel rows, the
// partial Gram matrices, the task's R, C and convergence state) issue
// together at entry.  Panel rows are loaded for the host's bounds on R and C
// (inside the task's workspace slot) and masked once the task's own R and C
// have arrived.
#include "tn_device.h"
#include "tn_devmath.h"
#include "tn_kernels.h"

#include <algorithm>
#include <cfloat>

namespace tnmps {

// Block at position i (1 .. nbe-1) of the circle at `step` (0 .. nbe-2):
// 1 + (i - 1 + step) mod (nbe - 1), the sum being below 2 (nbe - 1).
__device__ __forceinline__ int rr_player(int i, int step, int nbe) {
  int v = i - 1 + step;
  if (v >= nbe - 1) v -= nbe - 1;
  return 1 + v;
}

// Columns whose squared norm is below a.tiny2 (by default norm 1.5e-16,
// under f64::EPSILON times the unit-norm SVD input, so below every
// truncation threshold of the rank rule) take part in no rotation and in no
// convergence measure.

template <int BW>
struct PairSmem {
  static constexpr int PW = 2 * BW, LD = PW + 1;
  double2 Gs[2][PW][LD];            // panel Gram matrix (Gs[0]), double-buffered in the solve
  double2 buf[64 * LD + 2 * PW * LD];
  double red[4];
  int rr, go;  // round word and decision of the two-wave 16 x 16 solve
  // buf holds the 64-row panel chunk Ps and the accumulated rotation Ws
  // (double-buffered); the per-wave partial Gram matrices of panel_gram
  // overlay both once its row loop is done.
  static_assert(4 * PW * LD <= 64 * LD + 2 * PW * LD, "partial Gram overlay");
  __device__ double2& Ps(int r, int c) { return buf[r * LD + c]; }
  __device__ double2& Ws(int b, int i, int j) { return buf[64 * LD + (b * PW + i) * LD + j]; }
  __device__ double2& part(int w, int i, int j) { return buf[(w * PW + i) * LD + j]; }
};

// Column of X (and W) held by panel column cc (0..PW-1) of the pair of
// blocks (ba, bb).
template <int BW>
__device__ __forceinline__ int pair_col(int ba, int bb, int cc) {
  return ((cc < BW) ? ba : bb) * BW + (cc & (BW - 1));
}

// One 64-row chunk of a panel in registers: thread tid holds row tid & 63 of
// panel columns (tid >> 6) + 4 q, q < BW / 2.  Chunks are loaded one ahead
// of the LDS compute, so global-memory latency overlaps the work on the
// previous chunk (and, in the apply kernel, the panel solve).
template <int BW>
struct Chunk {
  double2 v[BW / 2];
};

// Columns of the thread's chunk elements (`lcol`) and of its outputs in
// panel_apply (`ocol`: for BW = 8 the matrix-core output column lane & 15 in
// ocol[0]; otherwise panel columns (tid >> 6) * BW / 2 + k).
template <int BW>
__device__ __forceinline__ void thread_cols(int nbe, int step, int p, int (&lcol)[BW / 2],
                                            int (&ocol)[BW / 2]) {
  const int wid = threadIdx.x >> 6;
  // pair p of `step` holds blocks ba and bb (the same for the whole workgroup)
  const int ba = (p == 0) ? 0 : rr_player(p, step, nbe);
  const int bb = rr_player(nbe - 1 - p, step, nbe);
#pragma unroll
  for (int q = 0; q < BW / 2; ++q) {
    lcol[q] = pair_col<BW>(ba, bb, wid + 4 * q);
    ocol[q] = (BW == 8) ? pair_col<BW>(ba, bb, threadIdx.x & 15)
                        : pair_col<BW>(ba, bb, wid * (BW / 2) + q);
  }
}

// Rows [r0, min(r0 + 64, rend)) of the panel of M, columns below cend.
template <int BW>
__device__ __forceinline__ void load_chunk(const double2* __restrict__ M, size_t ld, int r0,
                                           int rend, int cend, const int (&lcol)[BW / 2],
                                           Chunk<BW>& ck) {
  const int r = r0 + (threadIdx.x & 63);
#pragma unroll
  for (int q = 0; q < BW / 2; ++q)
    ck.v[q] = (lcol[q] < cend && r < rend) ? M[(size_t)lcol[q] * ld + r] : make_double2(0.0, 0.0);
}

// Zero the elements of a speculatively loaded chunk outside rows < rend and
// columns < cend.
template <int BW>
__device__ __forceinline__ void mask_chunk(int r0, int rend, int cend, const int (&lcol)[BW / 2],
                                           Chunk<BW>& ck) {
  const int r = r0 + (threadIdx.x & 63);
#pragma unroll
  for (int q = 0; q < BW / 2; ++q)
    if (!(lcol[q] < cend && r < rend)) ck.v[q] = make_double2(0.0, 0.0);
}

template <int BW>
__device__ __forceinline__ void put_chunk(const Chunk<BW>& ck, PairSmem<BW>& sm) {
  const int rr = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < BW / 2; ++q) sm.Ps(rr, wid + 4 * q) = ck.v[q];
}

// FP64 matrix-core tile D(16x16) += A(16x4) B(4x16): lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16]; D[i][j] is element i / 4 of lane
// j + 16 (i % 4) (lane l, element v: row l / 16 + 4 v, column l % 16).
typedef double v4d __attribute__((ext_vector_type(4)));
__device__ __forceinline__ v4d mfma16(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Panel Gram matrix G = P^H P of 16 columns on the matrix cores with three
// real products per row quad instead of four: Re G = Pr^T Pr + Pi^T Pi and
// M = Pr^T Pi, from which Im G = M - M^T (exactly antisymmetric, zero on
// the diagonal).  Lane l holds P[k][l % 16] of the quad's row k = l / 16.
__device__ __forceinline__ void gram8_acc(double2 p, v4d& re, v4d& m) {
  re = mfma16(p.x, p.x, re);
  m = mfma16(p.x, p.y, m);
  re = mfma16(p.y, p.y, re);
}

// The four waves' (Re G, M) parts summed in a fixed order into sm.Gs[0],
// with Im G = S - S^T for S the summed M.
template <typename Smem>
__device__ __forceinline__ void gram8_finish(v4d re, v4d m, Smem& sm) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
#pragma unroll
  for (int v = 0; v < 4; ++v)
    sm.part(wid, (lane >> 4) + 4 * v, lane & 15) = make_double2(re[v], m[v]);
  __syncthreads();
  {
    const int gi = tid >> 4, gj = tid & 15;
    double2 g = sm.part(0, gi, gj);
    double mt = sm.part(0, gj, gi).y;
#pragma unroll
    for (int w = 1; w < 4; ++w) {
      g = cadd(g, sm.part(w, gi, gj));
      mt += sm.part(w, gj, gi).y;
    }
    sm.Gs[0][gi][gj] = make_double2(g.x, g.y - mt);
  }
  __syncthreads();
}

// O += P Wl for one 16-row tile of 16 columns, three real products instead
// of four: T1 = Pr Wr, T2 = Pi Wi, T3 = (Pr + Pi)(Wr + Wi); Re O = T1 - T2,
// Im O = T3 - T1 - T2 (wks = Wr + Wi).  Lane l holds P[l % 16][4 kk + l / 16]
// in p[kk].
__device__ __forceinline__ void apply8_tile(const double2 (&p)[4], const double2 (&wk)[4],
                                            const double (&wks)[4], v4d& ore, v4d& oim) {
  v4d t1 = {0.0, 0.0, 0.0, 0.0}, t2 = {0.0, 0.0, 0.0, 0.0}, t3 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    t1 = mfma16(p[kk].x, wk[kk].x, t1);
    t2 = mfma16(p[kk].y, wk[kk].y, t2);
    t3 = mfma16(p[kk].x + p[kk].y, wks[kk], t3);
  }
  ore = t1 - t2;
  oim = t3 - t1 - t2;
}

// sm.Gs[0] = P^H P over rows [a, b) of the panel of X; `ck` holds rows
// [a, a + 64).  Thread = (row slice rs, 4x4 tile of G); a row slice takes
// every RS-th row of each 64-row chunk.  Slices are reduced by shuffles
// (inside a wave) and through LDS (across the four waves).
template <int BW>
__device__ void panel_gram(const double2* __restrict__ X, size_t ldx, int a, int b, int cend,
                           const int (&lcol)[BW / 2], Chunk<BW>& ck, PairSmem<BW>& sm) {
  if constexpr (BW == 8) {
    // 16-column panels on the matrix cores: wave w takes rows 16 w .. 16 w + 15
    // of each chunk, four at a time, as the K dimension (gram8_acc)
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    v4d gre = {0.0, 0.0, 0.0, 0.0}, gm = {0.0, 0.0, 0.0, 0.0};
    for (int r0 = a; r0 < b; r0 += 64) {
      put_chunk<BW>(ck, sm);
      __syncthreads();
      if (r0 + 64 < b) load_chunk<BW>(X, ldx, r0 + 64, b, cend, lcol, ck);
#pragma unroll
      for (int g = 0; g < 4; ++g)
        gram8_acc(sm.Ps(16 * wid + 4 * g + (lane >> 4), lane & 15), gre, gm);
      __syncthreads();
    }
    gram8_finish(gre, gm, sm);
    return;
  }
  constexpr int PW = 2 * BW, TPR = PW / 4, NTILE = TPR * TPR, RS = 256 / NTILE;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int tile = tid % NTILE, rs = tid / NTILE;
  const int ti = tile / TPR, tj = tile % TPR;
  double2 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = make_double2(0.0, 0.0);
  for (int r0 = a; r0 < b; r0 += 64) {
    put_chunk<BW>(ck, sm);
    __syncthreads();
    if (r0 + 64 < b) load_chunk<BW>(X, ldx, r0 + 64, b, cend, lcol, ck);
#pragma unroll 2
    for (int k = 0; k < 64 / RS; ++k) {
      const int rr = rs + RS * k;
      double2 xi[4], xj[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) xi[i] = sm.Ps(rr, ti * 4 + i);
#pragma unroll
      for (int j = 0; j < 4; ++j) xj[j] = sm.Ps(rr, tj * 4 + j);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = cadd(acc[i][j], cmulc(xi[i], xj[j]));
    }
    __syncthreads();
  }
  // lanes tile, tile + NTILE, ... of one wave hold the same tile
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double2 v = acc[i][j];
#pragma unroll
      for (int off = NTILE; off < 64; off <<= 1) {
        v.x += __shfl_xor(v.x, off, 64);
        v.y += __shfl_xor(v.y, off, 64);
      }
      acc[i][j] = v;
    }
  if (lane < NTILE) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) sm.part(wid, ti * 4 + i, tj * 4 + j) = acc[i][j];
  }
  __syncthreads();
  for (int e = tid; e < PW * PW; e += 256) {
    const int gi = e / PW, gj = e % PW;
    double2 g = sm.part(0, gi, gj);
    g = cadd(g, sm.part(1, gi, gj));
    g = cadd(g, sm.part(2, gi, gj));
    g = cadd(g, sm.part(3, gi, gj));
    sm.Gs[0][gi][gj] = g;
  }
  __syncthreads();
}

// 1/sqrt(x) from the hardware approximation (relative error about 5e-8 on
// gfx950) refined by one third-order Newton step,
// r (1 + e/2 + 3e^2/8) with e = 1 - x r^2: within 1.3 ulp, and four
// dependent operations on the rotation's critical path.
__device__ __forceinline__ double rsq_ref(double x) {
  const double r = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * r, r, 1.0);
  return fma(r * e, fma(e, 0.375, 0.5), r);
}

// Jacobi rotation zeroing the off-diagonal of [[a, c], [conj(c), b]]:
// Q = [[cs, sn], [-conj(sn), cs]], the rotation by the smaller angle
// (|theta| <= pi/4) with t = tan(theta) = 2|c| / s, s = |d| + r, d = b - a,
// r = sqrt(d^2 + 4|c|^2) (from the hardware reciprocal square root), written
// cs = s / sqrt(s^2 + 4|c|^2) and sn = 2 c / sqrt(s^2 + 4|c|^2) times the sign
// of d.  The angle carries the approximation's error (about 1e-7 relative:
// the pair keeps 1e-7 of its off-diagonal, which the next visit removes);
// unitarity does not, since cs^2 + |sn|^2 = (s^2 + 4|c|^2) / u with
// 1/sqrt(u) refined to about 1 ulp from the same u.  Identity (and false)
// when the pair is already orthogonal to sqrt(tol2) or a column is
// negligible; branch-free, the selects come last.
__device__ __forceinline__ bool rot2(double a, double b, double2 c, double tol2, double tiny2,
                                     double& cs, double2& sn) {
  const double c2 = cabs2(c);
  const bool go = a > tiny2 && b > tiny2 && c2 > tol2 * a * b;
  const double d = b - a;
  const double c4 = 4.0 * c2;
  const double r2 = fma(d, d, c4);
  const double s = fma(r2, __builtin_amdgcn_rsq(r2), fabs(d));
  const double ru = rsq_ref(fma(s, s, c4));
  const double f = (d >= 0.0 ? 2.0 : -2.0) * ru;
  cs = go ? s * ru : 1.0;
  sn = go ? make_double2(f * c.x, f * c.y) : make_double2(0.0, 0.0);
  return go;
}

// rot2 of two independent pairs, the two dependency chains interleaved
// statement by statement (the same arithmetic, bitwise the same results).
__device__ __forceinline__ void rot2x2(double a1, double b1, double2 c1, double a2, double b2,
                                       double2 c2, double tol2, double tiny2, double& cs1,
                                       double2& sn1, double& cs2, double2& sn2) {
  const double c21 = cabs2(c1), c22 = cabs2(c2);
  const bool go1 = a1 > tiny2 && b1 > tiny2 && c21 > tol2 * a1 * b1;
  const bool go2 = a2 > tiny2 && b2 > tiny2 && c22 > tol2 * a2 * b2;
  const double d1 = b1 - a1, d2 = b2 - a2;
  const double c41 = 4.0 * c21, c42 = 4.0 * c22;
  const double r21 = fma(d1, d1, c41), r22 = fma(d2, d2, c42);
  const double q1 = __builtin_amdgcn_rsq(r21), q2 = __builtin_amdgcn_rsq(r22);
  const double s1 = fma(r21, q1, fabs(d1)), s2 = fma(r22, q2, fabs(d2));
  const double u1 = fma(s1, s1, c41), u2 = fma(s2, s2, c42);
  const double h1 = __builtin_amdgcn_rsq(u1), h2 = __builtin_amdgcn_rsq(u2);
  const double e1 = fma(-u1 * h1, h1, 1.0), e2 = fma(-u2 * h2, h2, 1.0);
  const double ru1 = fma(h1 * e1, fma(e1, 0.375, 0.5), h1);
  const double ru2 = fma(h2 * e2, fma(e2, 0.375, 0.5), h2);
  const double f1 = (d1 >= 0.0 ? 2.0 : -2.0) * ru1, f2 = (d2 >= 0.0 ? 2.0 : -2.0) * ru2;
  cs1 = go1 ? s1 * ru1 : 1.0;
  cs2 = go2 ? s2 * ru2 : 1.0;
  sn1 = go1 ? make_double2(f1 * c1.x, f1 * c1.y) : make_double2(0.0, 0.0);
  sn2 = go2 ? make_double2(f2 * c2.x, f2 * c2.y) : make_double2(0.0, 0.0);
}

// Orders one wave's LDS accesses across rounds: LDS instructions of a wave
// are performed in issue order, so a compiler barrier is enough (no wait for
// the stores to complete).
__device__ __forceinline__ void wave_sync() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}

// ------------------------------------------------- 16 x 16 solve, two waves
// One cyclic two-sided Jacobi sweep of a 16 x 16 panel Gram matrix on two
// waves of the workgroup: wave 0 runs the rounds on G (G[0] -> G[1] -> ...
// double-buffered) and posts each round's 8 rotations to LDS; wave 1 follows
// round by round and applies them to W, in place from the identity.  Round r
// pairs index x with x ^ r (x with a zero bit inserted at r's highest set
// bit).  The LDS performs a wave's operations in issue order and serves the
// waves of a workgroup from one memory, so wave 0 orders its rounds with
// compiler barriers alone, and wave 1, once it reads round r on the round
// word (written after the round's rotations), reads those rotations.  Wave 0
// never waits, so wave 1's waits end (they are bounded as well).  The caller
// clears *rr before a barrier that precedes both waves' calls, and meets the
// two waves at a barrier afterwards.
constexpr int kRoundsNone = 100;    // posted when there is nothing to rotate
constexpr int kSolveSpin = 1 << 22;

struct Solve16 {
  double2 (*G)[16][17];  // G[0]: the panel Gram matrix
  double2 (*W)[17];      // 16 rows: the accumulated rotation
  double* rcs;           // [16][8]: cosine of round r, pair q at rcs[8 r + q]
  double2* rsn;          // [16][8]: sine
  int* rr;               // last round posted
};

// Wave 0, round ROUND: lane l = 8 g + i takes row pair g and column pair
// (i + g) % 8, so every group of 8 lanes covers the 8 column pairs and its
// first lane's column pair is its row pair.  The lane computes its column
// pair's rotation (the eight lanes of a pair compute bitwise the same one),
// takes its row pair's from the first lane of its group on the DPP network,
// computes its 2 x 2 block of G' = Qa^H G Qb and posts its column pair's
// rotation.  Reads G[CUR], writes G[CUR ^ 1].
template <int ROUND, int CUR, bool HOIST>
__device__ __forceinline__ void g_round16(const Solve16& sv, double tol2, double tiny2) {
  constexpr int NXT = CUR ^ 1;
  constexpr int h = ROUND >= 8 ? 3 : ROUND >= 4 ? 2 : ROUND >= 2 ? 1 : 0;
  constexpr int lowm = (1 << h) - 1;
  auto first = [](int q) { return ((q >> h) << (h + 1)) | (q & lowm); };
  // unless HOIST, the lane index passes through an empty asm so that the
  // round's LDS addresses are formed here, not hoisted out of the 15 rounds
  // (their registers cost the one-per-pair kernel occupancy; the row-split
  // apply kernel, with few workgroups, takes them and saves the address
  // arithmetic of every round)
  int lane = threadIdx.x & 63;
  if constexpr (!HOIST) asm volatile("" : "+v"(lane));
  double2 (*G)[17] = sv.G[CUR];
  double2 (*Gn)[17] = sv.G[NXT];
  const int pr = lane >> 3, pc = (lane + pr) & 7;
  const int x1 = first(pr), x2 = x1 ^ ROUND;
  const int y1 = first(pc), y2 = y1 ^ ROUND;
  const double ay = G[y1][y1].x, by = G[y2][y2].x;
  const double2 cy = G[y1][y2];
  const double2 g11 = G[x1][y1], g12 = G[x1][y2];
  const double2 g21 = G[x2][y1], g22 = G[x2][y2];
  // the round's loads all issue before the rotation's arithmetic
  __builtin_amdgcn_sched_barrier(0);
  double csb;
  double2 snb;
  rot2(ay, by, cy, tol2, tiny2, csb, snb);
  const double csa = bcast8_first(csb);
  const double2 sna = make_double2(bcast8_first(snb.x), bcast8_first(snb.y));
  // T = G[block] Qb, then G' = Qa^H T, with Q = [[cs, sn], [-conj(sn), cs]]
  const double2 msb = make_double2(-snb.x, snb.y);  // -conj(sn_b)
  const double2 t11 = cadd(cscale(g11, csb), cmul(g12, msb));
  const double2 t12 = cadd(cmul(g11, snb), cscale(g12, csb));
  const double2 t21 = cadd(cscale(g21, csb), cmul(g22, msb));
  const double2 t22 = cadd(cmul(g21, snb), cscale(g22, csb));
  const double2 nsa = make_double2(-sna.x, -sna.y);  // -sn_a
  const double2 csa_ = cconj(sna);                    // conj(sn_a)
  Gn[x1][y1] = cadd(cscale(t11, csa), cmul(nsa, t21));
  Gn[x1][y2] = cadd(cscale(t12, csa), cmul(nsa, t22));
  Gn[x2][y1] = cadd(cmul(csa_, t11), cscale(t21, csa));
  Gn[x2][y2] = cadd(cmul(csa_, t12), cscale(t22, csa));
  sv.rcs[8 * ROUND + pc] = csb;
  sv.rsn[8 * ROUND + pc] = snb;
}

template <int ROUND, int CUR, bool HOIST>
__device__ __forceinline__ void g_rounds16(const Solve16& sv, double tol2, double tiny2) {
  if constexpr (ROUND < 16) {
    g_round16<ROUND, CUR, HOIST>(sv, tol2, tiny2);
    wave_sync();
    __hip_atomic_store(sv.rr, ROUND, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    wave_sync();
    g_rounds16<ROUND + 1, CUR ^ 1, HOIST>(sv, tol2, tiny2);
  }
}

// Wave 1, round ROUND: lane l rotates rows l / 8 and l / 8 + 8 of W in
// column pair l % 8 (W' = W Qb), in place.
template <int ROUND, bool HOIST>
__device__ __forceinline__ void w_round16(const Solve16& sv) {
  constexpr int h = ROUND >= 8 ? 3 : ROUND >= 4 ? 2 : ROUND >= 2 ? 1 : 0;
  constexpr int lowm = (1 << h) - 1;
  auto first = [](int q) { return ((q >> h) << (h + 1)) | (q & lowm); };
  int lane = threadIdx.x & 63;
  if constexpr (!HOIST) asm volatile("" : "+v"(lane));
  const int q = lane & 7, y1 = first(q), y2 = y1 ^ ROUND;
  const int i0 = lane >> 3, i1 = i0 + 8;
  const double cs = sv.rcs[8 * ROUND + q];
  const double2 sn = sv.rsn[8 * ROUND + q];
  const double2 msn = make_double2(-sn.x, sn.y);  // -conj(sn)
  const double2 w01 = sv.W[i0][y1], w02 = sv.W[i0][y2];
  const double2 w11 = sv.W[i1][y1], w12 = sv.W[i1][y2];
  sv.W[i0][y1] = cadd(cscale(w01, cs), cmul(w02, msn));
  sv.W[i0][y2] = cadd(cmul(w01, sn), cscale(w02, cs));
  sv.W[i1][y1] = cadd(cscale(w11, cs), cmul(w12, msn));
  sv.W[i1][y2] = cadd(cmul(w11, sn), cscale(w12, cs));
}

__device__ __forceinline__ void wait_round(const int* rr, int r) {
  for (int it = 0; it < kSolveSpin; ++it) {
    if (__hip_atomic_load(rr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= r) break;
    __builtin_amdgcn_s_sleep(1);
  }
  wave_sync();
}

template <int ROUND, bool HOIST>
__device__ __forceinline__ void w_rounds16(const Solve16& sv) {
  if constexpr (ROUND < 16) {
    wait_round(sv.rr, ROUND);
    w_round16<ROUND, HOIST>(sv);
    wave_sync();
    w_rounds16<ROUND + 1, HOIST>(sv);
  }
}

// Wave 0's part: the largest relative off-diagonal of G[0] (recorded into
// *rec when not null and positive), then the rounds unless it is at most
// tol_rot: rounds 1 .. 15 when `inner`, else rounds 8 .. 15 only (the pairs
// across the two 8-column blocks; rounds 1 .. 7 pair columns inside a
// block).  Returns whether there is a rotation to apply.
template <bool HOIST>
__device__ bool solve16_wave0(const Solve16& sv, double* rec, double tol_rot, double tiny2,
                              bool inner) {
  const int lane = threadIdx.x & 63;
  double r2 = 0.0;
  // the four ratios straight-line (every division issued, masked after) and
  // their maximum over the wave on the DPP network
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = lane + 64 * q, gi = e >> 4, gj = e & 15;
    const double ga = sv.G[0][gi][gi].x, gb = sv.G[0][gj][gj].x;
    const bool ok = gi < gj && ga > tiny2 && gb > tiny2;
    const double rt = cabs2(sv.G[0][gi][gj]) / (ok ? ga * gb : 1.0);
    r2 = fmax(r2, ok ? rt : 0.0);
  }
  r2 = wave_max_dpp(r2);
  const double off = sqrt(r2);
  if (rec && lane == 0 && off > 0.0)
    atomicMax((unsigned long long*)rec, (unsigned long long)__double_as_longlong(off));
  if (!(off > tol_rot)) {
    __hip_atomic_store(sv.rr, kRoundsNone, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    return false;
  }
  if (inner)
    g_rounds16<1, 0, HOIST>(sv, tol_rot * tol_rot, tiny2);
  else
    g_rounds16<8, 0, HOIST>(sv, tol_rot * tol_rot, tiny2);
  return true;
}

// Wave 1's part: W from the identity through the posted rounds (from round
// 1 when `inner`, else from round 8, as wave 0).
template <bool HOIST>
__device__ void solve16_wave1(const Solve16& sv, bool inner) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = lane + 64 * q;
    sv.W[e >> 4][e & 15] = make_double2((e >> 4) == (e & 15) ? 1.0 : 0.0, 0.0);
  }
  wave_sync();
  wait_round(sv.rr, inner ? 1 : 8);
  if (__hip_atomic_load(sv.rr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == kRoundsNone)
    return;
  if (inner)
    w_rounds16<1, HOIST>(sv);
  else
    w_rounds16<8, HOIST>(sv);
}

// From the panel Gram matrix in sm.Gs[0]: record the largest relative
// off-diagonal into *rec (when not null), and unless the panel is already
// orthogonal apply one cyclic two-sided Jacobi sweep to G (G <- Q^H G Q),
// accumulating the rotation in sm.Ws[wsel].  Round r = 1 .. PW-1 pairs index
// x with x ^ r: PW/2 disjoint rotations per round, every pair once per sweep.
// 16 x 16 panels (BW = 8) take the two-wave solve above; wider ones spread a
// round's work items over the workgroup, each computing the rotations it
// needs from the round's Gram buffer, one barrier per round.  Returns false
// (for the whole block) when there is nothing to apply.  A function of
// sm.Gs[0] alone: blocks holding the same G compute bitwise the same
// rotation.  `inner` false leaves out the 16 x 16 solve's rounds inside each
// 8-column block (see solve16_wave0).
template <int BW, bool HOIST = false, bool RR_READY = false>
__device__ bool panel_solve(double* rec, double tol_rot, double tiny2, PairSmem<BW>& sm,
                            int& wsel, bool inner = true) {
  constexpr int PW = 2 * BW;
  const int tid = threadIdx.x;
  if constexpr (BW == 8) {
    // waves 0 and 1 (the panel chunk buffer holds the posted rotations: the
    // chunks are in registers meanwhile); W lands in Ws[0]
    const Solve16 sv{sm.Gs, (double2(*)[17]) & sm.Ws(0, 0, 0), (double*)&sm.Ps(0, 0),
                     &sm.Ps(8, 0), &sm.rr};
    // (RR_READY: the caller zeroed sm.rr, and a barrier after that and after
    // its last use of sm.Gs[0] and sm.buf precedes this call)
    if constexpr (!RR_READY) {
      if (tid == 0) sm.rr = 0;
      __syncthreads();
    }
    const int wid = tid >> 6;
    if (wid == 0) {
      const bool go = solve16_wave0<HOIST>(sv, rec, tol_rot, tiny2, inner);
      if (tid == 0) sm.go = go ? 1 : 0;
    } else if (wid == 1) {
      solve16_wave1<HOIST>(sv, inner);
    }
    __syncthreads();
    wsel = 0;
    return sm.go != 0;
  } else {
  const double tol2 = tol_rot * tol_rot;
  double ratio2 = 0.0;
  for (int e = tid; e < PW * PW; e += 256) {
    const int gi = e / PW, gj = e % PW;
    sm.Ws(0, gi, gj) = make_double2(gi == gj ? 1.0 : 0.0, 0.0);
    if (gi < gj) {
      const double a = sm.Gs[0][gi][gi].x, b = sm.Gs[0][gj][gj].x;
      if (a > tiny2 && b > tiny2) ratio2 = fmax(ratio2, cabs2(sm.Gs[0][gi][gj]) / (a * b));
    }
  }
  const double off = sqrt(block_max(ratio2, sm.red));
  if (rec && tid == 0 && off > 0.0)
    atomicMax((unsigned long long*)rec, (unsigned long long)__double_as_longlong(off));
  if (off <= tol_rot) return false;
  // Work items of a round: the (PW/2)^2 2x2 blocks of G on the round's row
  // pair x column pair (G' = Qa^H G Qb touches only its own block), then the
  // PW x PW/2 (row, column pair) items of W' = W Qb.  Pair a of round r is
  // (x, x ^ r) with x = a with a zero bit inserted at r's highest set bit h.
  constexpr int HP = PW / 2, NG = HP * HP, NW = PW * HP;
  int cur = 0;
  for (int round = 1; round < PW; ++round) {
    const int h = 31 - __builtin_clz(round);
    const int lowm = (1 << h) - 1;
    for (int e = tid; e < NG + NW; e += 256) {
      if (e < NG) {
        const int pa = e / HP, pb = e % HP;
        const int x1 = ((pa >> h) << (h + 1)) | (pa & lowm), x2 = x1 ^ round;
        const int y1 = ((pb >> h) << (h + 1)) | (pb & lowm), y2 = y1 ^ round;
        const double ax = sm.Gs[cur][x1][x1].x, bx = sm.Gs[cur][x2][x2].x;
        const double2 cx = sm.Gs[cur][x1][x2];
        const double ay = sm.Gs[cur][y1][y1].x, by = sm.Gs[cur][y2][y2].x;
        const double2 cy = sm.Gs[cur][y1][y2];
        const double2 g11 = sm.Gs[cur][x1][y1], g12 = sm.Gs[cur][x1][y2];
        const double2 g21 = sm.Gs[cur][x2][y1], g22 = sm.Gs[cur][x2][y2];
        double csa, csb;
        double2 sna, snb;
        rot2(ax, bx, cx, tol2, tiny2, csa, sna);
        rot2(ay, by, cy, tol2, tiny2, csb, snb);
        // T = G[block] Qb, then G' = Qa^H T, with Q = [[cs, sn], [-conj(sn), cs]]
        const double2 msb = make_double2(-snb.x, snb.y);  // -conj(sn_b)
        const double2 t11 = cadd(cscale(g11, csb), cmul(g12, msb));
        const double2 t12 = cadd(cmul(g11, snb), cscale(g12, csb));
        const double2 t21 = cadd(cscale(g21, csb), cmul(g22, msb));
        const double2 t22 = cadd(cmul(g21, snb), cscale(g22, csb));
        const double2 nsa = make_double2(-sna.x, -sna.y);  // -sn_a
        const double2 csa_ = cconj(sna);                    // conj(sn_a)
        sm.Gs[cur ^ 1][x1][y1] = cadd(cscale(t11, csa), cmul(nsa, t21));
        sm.Gs[cur ^ 1][x1][y2] = cadd(cscale(t12, csa), cmul(nsa, t22));
        sm.Gs[cur ^ 1][x2][y1] = cadd(cmul(csa_, t11), cscale(t21, csa));
        sm.Gs[cur ^ 1][x2][y2] = cadd(cmul(csa_, t12), cscale(t22, csa));
      } else {
        const int w = e - NG;
        const int i = w / HP, pb = w % HP;
        const int y1 = ((pb >> h) << (h + 1)) | (pb & lowm), y2 = y1 ^ round;
        const double ay = sm.Gs[cur][y1][y1].x, by = sm.Gs[cur][y2][y2].x;
        const double2 cy = sm.Gs[cur][y1][y2];
        const double2 w1 = sm.Ws(cur, i, y1), w2 = sm.Ws(cur, i, y2);
        double csb;
        double2 snb;
        rot2(ay, by, cy, tol2, tiny2, csb, snb);
        const double2 msb = make_double2(-snb.x, snb.y);
        sm.Ws(cur ^ 1, i, y1) = cadd(cscale(w1, csb), cmul(w2, msb));
        sm.Ws(cur ^ 1, i, y2) = cadd(cmul(w1, snb), cscale(w2, csb));
      }
    }
    cur ^= 1;
    __syncthreads();
  }
  wsel = cur;
  return true;
  }
}

// Rows [a, b) of the panel of M (leading dimension ld, columns below cend)
// <- panel * Ws[wsel].  Thread = (row lane 0..63, output group of PW/4 columns);
// rows are staged through LDS 64 at a time.  `ck` holds rows [a, a + 64).
template <int BW>
__device__ void panel_apply(double2* M, size_t ld, int a, int b, int cend,
                            const int (&lcol)[BW / 2], const int (&ocol)[BW / 2], Chunk<BW>& ck,
                            PairSmem<BW>& sm, int wsel) {
  if constexpr (BW == 8) {
    // 16-column panels on the matrix cores: wave w computes rows 16 w ..
    // 16 w + 15 of each chunk, O = P Wl with the panel columns as K:
    //   Re O = Pr Wr - Pi Wi,   Im O = Pr Wi + Pi Wr.
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    double2 wk[4];
    double wks[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      wk[kk] = sm.Ws(wsel, 4 * kk + (lane >> 4), lane & 15);
      wks[kk] = wk[kk].x + wk[kk].y;
    }
    const int oc = ocol[0];
    for (int r0 = a; r0 < b; r0 += 64) {
      put_chunk<BW>(ck, sm);
      __syncthreads();
      if (r0 + 64 < b) load_chunk<BW>(M, ld, r0 + 64, b, cend, lcol, ck);
      v4d ore, oim;
      double2 pk[4];
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) pk[kk] = sm.Ps(16 * wid + (lane & 15), 4 * kk + (lane >> 4));
      apply8_tile(pk, wk, wks, ore, oim);
      if (oc < cend) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = r0 + 16 * wid + (lane >> 4) + 4 * v;
          if (r < b) M[(size_t)oc * ld + r] = make_double2(ore[v], oim[v]);
        }
      }
      __syncthreads();
    }
    return;
  }
  constexpr int PW = 2 * BW, NJ = PW / 4;
  const int tid = threadIdx.x;
  const int rl = tid & 63, jg = tid >> 6;
  for (int r0 = a; r0 < b; r0 += 64) {
    put_chunk<BW>(ck, sm);
    __syncthreads();
    if (r0 + 64 < b) load_chunk<BW>(M, ld, r0 + 64, b, cend, lcol, ck);
    const int r = r0 + rl;
    double2 o[NJ];
#pragma unroll
    for (int k = 0; k < NJ; ++k) o[k] = make_double2(0.0, 0.0);
#pragma unroll 4
    for (int c = 0; c < PW; ++c) {
      const double2 x = sm.Ps(rl, c);
#pragma unroll
      for (int k = 0; k < NJ; ++k) o[k] = cadd(o[k], cmul(x, sm.Ws(wsel, c, jg * NJ + k)));
    }
    if (r < b) {
#pragma unroll
      for (int k = 0; k < NJ; ++k)
        if (ocol[k] < cend) M[(size_t)ocol[k] * ld + r] = o[k];
    }
    __syncthreads();
  }
}

// ------------------------------------------- register-resident panels (BW = 8)
// The 16-column panel rows a workgroup works on, held in registers for the
// whole step: chunk c of thread tid is row a + 64 c + (tid & 63) of panel
// columns (tid >> 6) + 4 q, q < 4 (the Chunk layout).  Chunk 0 loads at
// kernel entry and the other chunks as soon as the task is known to be
// unconverged, so a step waits on global memory about once instead of once
// per chunk, and the Gram matrix and the rotation read X once.  The
// arithmetic and its order are those of panel_gram / panel_apply.
template <int NC>
struct Panel8 {
  double2 v[NC][4];
};

// Chunks C0 .. C1-1 of rows [a, rend), columns below cend.
template <int NC, int C0, int C1>
__device__ __forceinline__ void load_panel8(const double2* __restrict__ M, size_t ld, int a,
                                            int rend, int cend, const int (&lcol)[4],
                                            Panel8<NC>& pn) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int c = C0; c < C1; ++c) {
    const int r = a + 64 * c + lane;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      pn.v[c][q] =
          (lcol[q] < cend && r < rend) ? M[(size_t)lcol[q] * ld + r] : make_double2(0.0, 0.0);
  }
}

// Zero chunk 0's elements outside rows [a, rend) and columns below cend
// (it was loaded for the host's bounds, before the task's own arrived).
template <int NC>
__device__ __forceinline__ void mask_panel8(int a, int rend, int cend, const int (&lcol)[4],
                                            Panel8<NC>& pn) {
  const int r = a + (threadIdx.x & 63);
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (!(lcol[q] < cend && r < rend)) pn.v[0][q] = make_double2(0.0, 0.0);
}

// Zero every chunk's elements outside rows [a, rend) and columns below cend.
template <int NC>
__device__ __forceinline__ void mask_panel8_all(int a, int rend, int cend, const int (&lcol)[4],
                                                Panel8<NC>& pn) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int r = a + 64 * c + lane;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (!(lcol[q] < cend && r < rend)) pn.v[c][q] = make_double2(0.0, 0.0);
  }
}

// Rows [lo, b) of pn's columns (below cend) written to M unchanged (pn's
// chunk c holds rows 64 c).
template <int NC>
__device__ __forceinline__ void panel_copy8r(double2* M, size_t ld, int lo, int b, int cend,
                                             const int (&lcol)[4], const Panel8<NC>& pn) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (64 * c + 64 <= lo) continue;
    const int r = 64 * c + lane;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (lcol[q] < cend && r >= lo && r < b) M[(size_t)lcol[q] * ld + r] = pn.v[c][q];
  }
}

template <int NC>
__device__ __forceinline__ void put_panel8(const Panel8<NC>& pn, int c, PairSmem<8>& sm) {
  const int rr = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 4; ++q) sm.Ps(rr, wid + 4 * q) = pn.v[c][q];
}

// sm.Gs[0] = P^H P over rows [a, b) of the panel held in pn (b - a <= 64 NC).
template <int NC>
__device__ void panel_gram8r(int a, int b, const Panel8<NC>& pn, PairSmem<8>& sm) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  // two accumulators per part (even and odd row quads), so that consecutive
  // matrix-core operations do not wait on each other
  v4d gre[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
  v4d gm[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (a + 64 * c >= b) break;
    put_panel8<NC>(pn, c, sm);
    __syncthreads();
    double2 pv[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) pv[g] = sm.Ps(16 * wid + 4 * g + (lane >> 4), lane & 15);
#pragma unroll
    for (int g = 0; g < 4; ++g) gram8_acc(pv[g], gre[g & 1], gm[g & 1]);
    __syncthreads();
  }
  gram8_finish(gre[0] + gre[1], gm[0] + gm[1], sm);
}

// Rows [a + 64 c0, b) of the panel of M (columns below cend) <- pn * Ws[wsel]
// (pn's chunk c holds rows a + 64 c).
template <int NC>
__device__ void panel_apply8r(double2* M, size_t ld, int a, int b, int cend, int oc,
                              const Panel8<NC>& pn, PairSmem<8>& sm, int wsel, int c0 = 0) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  double2 wk[4];
  double wks[4];
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    wk[kk] = sm.Ws(wsel, 4 * kk + (lane >> 4), lane & 15);
    wks[kk] = wk[kk].x + wk[kk].y;
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int r0 = a + 64 * c;
    if (r0 >= b) break;
    if (c < c0) continue;
    put_panel8<NC>(pn, c, sm);
    __syncthreads();
    v4d ore, oim;
    double2 pk[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) pk[kk] = sm.Ps(16 * wid + (lane & 15), 4 * kk + (lane >> 4));
    apply8_tile(pk, wk, wks, ore, oim);
    if (oc < cend) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = r0 + 16 * wid + (lane >> 4) + 4 * v;
        if (r < b) M[(size_t)oc * ld + r] = make_double2(ore[v], oim[v]);
      }
    }
    __syncthreads();
  }
}

// The whole panel of X (chunks 0 .. NCG-1) and the workgroup's slice of W
// (chunks NCG ..), transposed into LDS once, behind one barrier, for the
// Gram matrix and both rotations of a row-split step.  Chunk k row r column
// c at v[k][r][c] (rows padded to 17 entries, as Ps).
template <int NCG, int NCW>
struct PanelLds {
  double2 v[NCG + NCW][64][17];
};

template <int NC>
__device__ __forceinline__ void put_panels8(const Panel8<NC>& pn, double2 (*dst)[64][17]) {
  const int rr = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[c][rr][wid + 4 * q] = pn.v[c][q];
}

// sm.Gs[0] = P^H P over rows [0, b) of the panel in chunks[0 ..]: the
// matrix-core products of panel_gram8r in the same order into the same
// accumulators (bitwise the same G), with no barrier between chunks.
template <int NC>
__device__ void panel_gram8_lds(int b, const double2 (*chunks)[64][17], PairSmem<8>& sm) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  v4d gre[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
  v4d gm[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (64 * c >= b) break;
    double2 pv[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) pv[g] = chunks[c][16 * wid + 4 * g + (lane >> 4)][lane & 15];
#pragma unroll
    for (int g = 0; g < 4; ++g) gram8_acc(pv[g], gre[g & 1], gm[g & 1]);
  }
  gram8_finish(gre[0] + gre[1], gm[0] + gm[1], sm);
}

// Slice rows [xa, xe) of X's panel column oc (into Xo) and [wa, we) of W's
// <- (copies in LDS) * Wl, in 16-row tiles spread over the four waves (tile
// i on wave i % 4: the slice of X first, then the slice of W), each tile
// with panel_apply8r's products.  xa is a multiple of 16 (X's tiles lie in
// one 64-row chunk each); chunk c of W's copies holds rows wa + 64 c.
template <int NCG>
__device__ __forceinline__ void apply8_tiles_lds(double2* Xo, size_t ldx, int xa, int xe,
                                                 double2* W, size_t ldw, int wa, int we,
                                                 int cend, int oc,
                                                 const double2 (*chunks)[64][17],
                                                 const double2 (&wk)[4], const double (&wks)[4]) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nxt = xe > xa ? (xe - xa + 15) >> 4 : 0;
  const int nwt = we > wa ? (we - wa + 15) >> 4 : 0;
  for (int i = wid; i < nxt + nwt; i += 4) {
    const bool isx = i < nxt;
    // tile rows: X's absolute, W's relative to wa
    const int r0 = isx ? xa + 16 * i : 16 * (i - nxt);
    const double2(*ck)[17] = chunks[(isx ? 0 : NCG) + (r0 >> 6)];
    const int rb = r0 & 63;
    double2 pk[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) pk[kk] = ck[rb + (lane & 15)][4 * kk + (lane >> 4)];
    v4d ore, oim;
    apply8_tile(pk, wk, wks, ore, oim);
    double2* M = isx ? Xo : W;
    const size_t ld = isx ? ldx : ldw;
    const int base = isx ? r0 : wa + r0, b = isx ? xe : we;
    if (oc < cend) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = base + (lane >> 4) + 4 * v;
        if (r < b) M[(size_t)oc * ld + r] = make_double2(ore[v], oim[v]);
      }
    }
  }
}

// Row range of slice q of S (a power of two) over n rows (slices are
// multiples of 64 rows).
__device__ __forceinline__ void split_rows(int n, int S, int q, int& a, int& b) {
  const int sl = __builtin_ctz(S);
  const int per = (((n + S - 1) >> sl) + 63) / 64 * 64;
  a = min(n, q * per);
  b = min(n, a + per);
}

// The same with slices of a multiple of 16 rows (the one-launch kernel's).
__device__ __forceinline__ void split_rows16(int n, int S, int q, int& a, int& b) {
  const int sl = __builtin_ctz(S);
  const int per = (((n + S - 1) >> sl) + 15) / 16 * 16;
  a = min(n, q * per);
  b = min(n, a + per);
}

// Block pair p and row slice q of a row-split workgroup (grid npairs x S
// workgroups): block p + npairs q.  Workgroups are dealt to the eight XCDs
// in turn, so for npairs a multiple of 8 every workgroup of a pair runs on
// one XCD, and their loads of the pair's whole panel (and of its partial
// Gram matrices) share one L2.
__device__ __forceinline__ void split_block(int S, int& p, int& q) {
  const int npairs = (int)gridDim.x >> __builtin_ctz(S);
  p = (int)blockIdx.x % npairs;
  q = (int)blockIdx.x / npairs;
}

// Convergence state of task t before `sweep`: true when its previous sweep
// stayed at or below conv_tol.  The load is unconditional (row 0 at sweep
// 0, where it is not used), so no branch separates it from the kernel's
// other entry loads.
// Whether step's 16 x 16 panel solves rotate the pairs inside each 8-column
// block too (the pairs across the two blocks on every step): every k-th step
// of a sweep, k = inner_every but at most a quarter of the sweep's nbe - 1
// steps, so every sweep keeps at least four such steps.  Every block takes
// part in every step, so a block's own columns are rotated against each
// other every k steps.
__device__ __forceinline__ bool inner_step(const JacobiArgs& a, int step) {
  const int k = min(a.inner_every, max(1, (a.nbe - 1) >> 2));
  return step % k == 0;
}

__device__ __forceinline__ bool converged(const JacobiArgs& a, int t, int sweep) {
  const double o = a.offs[(size_t)(sweep > 0 ? sweep - 1 : 0) * a.nt + t];
  return sweep > 0 && o <= a.conv_tol;
}

// Keeps the compiler from moving later memory reads above this point: the
// register-resident kernels issue their panel loads, then read the task's
// R, C and convergence state, whose wait then does not hold the panel loads
// back.
__device__ __forceinline__ void issue_fence() { asm volatile("" ::: "memory"); }

// Block pair p of task t at one step: Gram matrix, solve, apply to X and W.
template <int BW>
__device__ __forceinline__ void step_task(const JacobiArgs& a, int t, int p, int sweep, int step,
                                          PairSmem<BW>& sm) {
  const int* inf = a.info + (size_t)t * kInfoLen;
  const int R = inf[kInfoRJ], C = inf[kInfoC];
  const bool done = converged(a, t, sweep);
  int lcol[BW / 2], ocol[BW / 2];
  thread_cols<BW>(a.nbe, step, p, lcol, ocol);
  double2* X = a.X + (size_t)t * a.xslot;
  double2* W = a.W + (size_t)t * a.wslot;
  Chunk<BW> ck;
  load_chunk<BW>(X, (size_t)a.ldx, 0, a.Rh, a.ldw, lcol, ck);
  if (done) return;
  mask_chunk<BW>(0, R, C, lcol, ck);
  panel_gram<BW>(X, (size_t)a.ldx, 0, R, C, lcol, ck, sm);
  Chunk<BW> cx, cw;
  load_chunk<BW>(X, (size_t)a.ldx, 0, R, C, lcol, cx);
  load_chunk<BW>(W, (size_t)a.ldw, 0, C, C, lcol, cw);
  int wsel = 0;
  if (!panel_solve<BW>(a.offs + (size_t)sweep * a.nt + t, a.tol_rot, a.tiny2, sm, wsel,
                       inner_step(a, step)))
    return;
  panel_apply<BW>(X, (size_t)a.ldx, 0, R, C, lcol, ocol, cx, sm, wsel);
  panel_apply<BW>(W, (size_t)a.ldw, 0, C, C, lcol, ocol, cw, sm, wsel);
}

// One launch per step, one workgroup per block pair: of task blockIdx.y, or,
// with an active-task list, of listed task blockIdx.y (the grid has at least
// as many rows as the list has tasks; converged tasks then cost no
// workgroup).  Both words of the list load together.
template <int BW>
__global__ __launch_bounds__(256, BW == 8 ? 2 : 1) void k_jacobi_step(JacobiArgs a, int sweep,
                                                                      int step) {
  __shared__ PairSmem<BW> sm;
  const int i = blockIdx.y;
  const int cnt = a.act ? a.act[0] : a.nt;
  const int t = a.act ? a.act[1 + i] : i;
  if (i >= cnt) return;
  step_task<BW>(a, t, blockIdx.x, sweep, step, sm);
}

// Active-task list of a sweep: act[0] = number of tasks not converged before
// `sweep` (all of them at sweep 0), act[1 ..] their indices in ascending
// order.  One workgroup of 1024 threads; thread i counts a run of
// ceil(nt / 1024) consecutive tasks, a scan gives each run its offset.
__global__ __launch_bounds__(1024) void k_active(JacobiArgs a, int sweep, int* act) {
  __shared__ int wsum[16];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int per = (a.nt + 1023) / 1024;
  const int t0 = min(a.nt, tid * per), t1 = min(a.nt, t0 + per);
  int c = 0;
  for (int t = t0; t < t1; ++t) c += converged(a, t, sweep) ? 0 : 1;
  int x = c;  // inclusive scan over the wave
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  int base = 0;
  for (int w = 0; w < wid; ++w) base += wsum[w];
  int pos = base + x - c;
  for (int t = t0; t < t1; ++t)
    if (!converged(a, t, sweep)) act[1 + pos++] = t;
  if (tid == 1023) act[0] = base + x;
}

// Row-split step, part 1: workgroup (p, q) writes the Gram matrix of row
// slice q of pair p's panel to gpart[p][q].
template <int BW>
__global__ __launch_bounds__(256, BW == 8 ? 2 : 1) void k_jacobi_gram(JacobiArgs a, int sweep,
                                                                      int step) {
  constexpr int PW = 2 * BW;
  __shared__ PairSmem<BW> sm;
  const int t = blockIdx.y;
  int p, q;
  split_block(a.S, p, q);
  const int* inf = a.info + (size_t)t * kInfoLen;
  const int R = inf[kInfoRJ], C = inf[kInfoC];
  const bool done = converged(a, t, sweep);
  int lcol[BW / 2], ocol[BW / 2];
  thread_cols<BW>(a.nbe, step, p, lcol, ocol);
  int xa, xb;
  split_rows(a.Rh, a.S, q, xa, xb);
  const double2* X = a.X + (size_t)t * a.xslot;
  Chunk<BW> ck;
  load_chunk<BW>(X, (size_t)a.ldx, xa, xb, a.ldw, lcol, ck);
  if (done) return;
  const int xe = min(xb, R);
  mask_chunk<BW>(xa, xe, C, lcol, ck);
  panel_gram<BW>(X, (size_t)a.ldx, xa, xe, C, lcol, ck, sm);
  double2* gp = a.gpart + (size_t)t * a.gslot + ((size_t)p * a.S + q) * (PW * PW);
  for (int e = threadIdx.x; e < PW * PW; e += 256) gp[e] = sm.Gs[0][e / PW][e % PW];
}

// Row-split step, part 2: every workgroup of pair p sums the S partial Gram
// matrices in the same order, so all of them solve the same G into the same
// Wl, and applies it to its own slice of the rows of X and of W.
template <int BW>
__global__ __launch_bounds__(256, BW == 8 ? 2 : 1) void k_jacobi_apply(JacobiArgs a, int sweep,
                                                                       int step) {
  constexpr int PW = 2 * BW, NE = (PW * PW + 255) / 256;
  __shared__ PairSmem<BW> sm;
  const int t = blockIdx.y;
  int p, q;
  split_block(a.S, p, q);
  const int* inf = a.info + (size_t)t * kInfoLen;
  const int R = inf[kInfoRJ], C = inf[kInfoC];
  const bool done = converged(a, t, sweep);
  int lcol[BW / 2], ocol[BW / 2];
  thread_cols<BW>(a.nbe, step, p, lcol, ocol);
  int xa, xb, wa, wb;
  split_rows(a.Rh, a.S, q, xa, xb);
  split_rows(a.Ch, a.S, q, wa, wb);
  double2* X = a.X + (size_t)t * a.xslot;
  double2* W = a.W + (size_t)t * a.wslot;
  const double2* gp = a.gpart + (size_t)t * a.gslot + (size_t)p * a.S * (PW * PW);
  // the slice's first chunks and the partial Gram matrices load together
  Chunk<BW> cx, cw;
  load_chunk<BW>(X, (size_t)a.ldx, xa, xb, a.ldw, lcol, cx);
  load_chunk<BW>(W, (size_t)a.ldw, wa, wb, a.ldw, lcol, cw);
  double2 g[NE][kMaxSplit];
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    const int e = threadIdx.x + 256 * k;
#pragma unroll
    for (int s = 0; s < kMaxSplit; ++s)
      g[k][s] = (s < a.S && e < PW * PW) ? gp[(size_t)s * (PW * PW) + e] : make_double2(0.0, 0.0);
  }
  if (done) return;
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    const int e = threadIdx.x + 256 * k;
    double2 sum = g[k][0];
#pragma unroll
    for (int s = 1; s < kMaxSplit; ++s) sum = cadd(sum, g[k][s]);
    if (e < PW * PW) sm.Gs[0][e / PW][e % PW] = sum;
  }
  __syncthreads();
  const int xe = min(xb, R), we = min(wb, C);
  mask_chunk<BW>(xa, xe, C, lcol, cx);
  mask_chunk<BW>(wa, we, C, lcol, cw);
  int wsel = 0;
  if (!panel_solve<BW>(q == 0 ? a.offs + (size_t)sweep * a.nt + t : nullptr, a.tol_rot,
                       a.tiny2, sm, wsel, inner_step(a, step)))
    return;
  panel_apply<BW>(X, (size_t)a.ldx, xa, xe, C, lcol, ocol, cx, sm, wsel);
  panel_apply<BW>(W, (size_t)a.ldw, wa, we, C, lcol, ocol, cw, sm, wsel);
}

// k_jacobi_step for BW = 8 with X's panel (R <= 64 NCX rows) in registers;
// W's panel streams through panel_apply (its chunk 0 loads during the
// solve).  The same arithmetic as k_jacobi_step<8>.  Holding W's panel too
// cost a wave per SIMD of occupancy and measured slower in batched runs.
template <int NCX>
__global__ __launch_bounds__(256, 2) void k_jacobi_step_r(JacobiArgs a, int sweep, int step) {
  __shared__ PairSmem<8> sm;
  const int i = blockIdx.y;
  const int cnt = a.act ? a.act[0] : a.nt;
  const int t = a.act ? a.act[1 + i] : i;
  if (i >= cnt) return;
  int lcol[4], ocol[4];
  thread_cols<8>(a.nbe, step, blockIdx.x, lcol, ocol);
  double2* X = a.X + (size_t)t * a.xslot;
  double2* W = a.W + (size_t)t * a.wslot;
  Panel8<NCX> px;
  load_panel8<NCX, 0, 1>(X, (size_t)a.ldx, 0, a.Rh, a.ldw, lcol, px);
  issue_fence();
  const int* inf = a.info + (size_t)t * kInfoLen;
  const int R = inf[kInfoRJ], C = inf[kInfoC];
  if (converged(a, t, sweep)) return;
  load_panel8<NCX, 1, NCX>(X, (size_t)a.ldx, 0, R, C, lcol, px);
  mask_panel8<NCX>(0, R, C, lcol, px);
  panel_gram8r<NCX>(0, R, px, sm);
  Chunk<8> cw;
  load_chunk<8>(W, (size_t)a.ldw, 0, C, C, lcol, cw);
  int wsel = 0;
  if (!panel_solve<8>(a.offs + (size_t)sweep * a.nt + t, a.tol_rot, a.tiny2, sm, wsel,
                      inner_step(a, step)))
    return;
  panel_apply8r<NCX>(X, (size_t)a.ldx, 0, R, C, ocol[0], px, sm, wsel);
  panel_apply<8>(W, (size_t)a.ldw, 0, C, C, lcol, ocol, cw, sm, wsel);
}

// k_jacobi_gram for BW = 8 with the slice's rows (at most 64 NC) in registers.
template <int NC>
__global__ __launch_bounds__(256, 2) void k_jacobi_gram_r(JacobiArgs a, int sweep, int step) {
  constexpr int PW = 16;
  __shared__ PairSmem<8> sm;
  const int t = blockIdx.y;
  int p, q;
  split_block(a.S, p, q);
  int lcol[4], ocol[4];
  thread_cols<8>(a.nbe, step, p, lcol, ocol);
  int xa, xb;
  split_rows(a.Rh, a.S, q, xa, xb);
  const double2* X = a.X + (size_t)t * a.xslot;
  Panel8<NC> px;
  load_panel8<NC, 0, 1>(X, (size_t)a.ldx, xa, xb, a.ldw, lcol, px);
  issue_fence();
  const int* inf = a.info + (size_t)t * kInfoLen;
  const int R = inf[kInfoRJ], C = inf[kInfoC];
  if (converged(a, t, sweep)) return;
  const int xe = min(xb, R);
  load_panel8<NC, 1, NC>(X, (size_t)a.ldx, xa, xe, C, lcol, px);
  mask_panel8<NC>(xa, xe, C, lcol, px);
  panel_gram8r<NC>(xa, xe, px, sm);
  double2* gp = a.gpart + (size_t)t * a.gslot + ((size_t)p * a.S + q) * (PW * PW);
  for (int e = threadIdx.x; e < PW * PW; e += 256) gp[e] = sm.Gs[0][e / PW][e % PW];
}

// k_jacobi_apply for BW = 8 with the slice's rows of X and of W (at most
// 64 NC each) in registers.
template <int NC>
__global__ __launch_bounds__(256, 2) void k_jacobi_apply_r(JacobiArgs a, int sweep, int step) {
  constexpr int PW = 16;
  __shared__ PairSmem<8> sm;
  const int t = blockIdx.y;
  int p, q;
  split_block(a.S, p, q);
  int lcol[4], ocol[4];
  thread_cols<8>(a.nbe, step, p, lcol, ocol);
  int xa, xb, wa, wb;
  split_rows(a.Rh, a.S, q, xa, xb);
  split_rows(a.Ch, a.S, q, wa, wb);
  double2* X = a.X + (size_t)t * a.xslot;
  double2* W = a.W + (size_t)t * a.wslot;
  const double2* gp = a.gpart + (size_t)t * a.gslot + (size_t)p * a.S * (PW * PW);
  Panel8<NC> px, pw;
  load_panel8<NC, 0, 1>(X, (size_t)a.ldx, xa, xb, a.ldw, lcol, px);
  load_panel8<NC, 0, 1>(W, (size_t)a.ldw, wa, wb, a.ldw, lcol, pw);
  double2 g[kMaxSplit];
#pragma unroll
  for (int s = 0; s < kMaxSplit; ++s)
    g[s] = s < a.S ? gp[(size_t)s * (PW * PW) + threadIdx.x] : make_double2(0.0, 0.0);
  issue_fence();
  const int* inf = a.info + (size_t)t * kInfoLen;
  const int R = inf[kInfoRJ], C = inf[kInfoC];
  if (converged(a, t, sweep)) return;
  const int xe = min(xb, R), we = min(wb, C);
  load_panel8<NC, 1, NC>(X, (size_t)a.ldx, xa, xe, C, lcol, px);
  load_panel8<NC, 1, NC>(W, (size_t)a.ldw, wa, we, C, lcol, pw);
  {
    double2 sum = g[0];
#pragma unroll
    for (int s = 1; s < kMaxSplit; ++s) sum = cadd(sum, g[s]);
    sm.Gs[0][threadIdx.x / PW][threadIdx.x % PW] = sum;
  }
  __syncthreads();
  mask_panel8<NC>(xa, xe, C, lcol, px);
  mask_panel8<NC>(wa, we, C, lcol, pw);
  int wsel = 0;
  if (!panel_solve<8, true>(q == 0 ? a.offs + (size_t)sweep * a.nt + t : nullptr, a.tol_rot,
                            a.tiny2, sm, wsel, inner_step(a, step)))
    return;
  panel_apply8r<NC>(X, (size_t)a.ldx, xa, xe, C, ocol[0], px, sm, wsel);
  panel_apply8r<NC>(W, (size_t)a.ldw, wa, we, C, ocol[0], pw, sm, wsel);
}

// Row-split step in one launch: workgroup (p, q) forms the Gram matrix of
// pair p's whole panel (R <= 64 NCG rows; every workgroup of the pair forms
// the same G in the same order, as k_jacobi_step_r does), solves it, and
// writes its slice q of the rotated rows of X and of W (slices of a
// multiple of 16 rows, W's at most 64 NCW rows).  The panel and the slice
// of W are loaded into registers at entry and transposed into LDS once,
// behind one barrier: the Gram matrix's matrix-core products run back to
// back from there, and the rotation's 16-row tiles of both slices, spread
// over the four waves, read the same copies.  PairSmem and PanelLds take
// about 136 KiB of LDS (NCW = 2; 119 KiB for NCW = 1), one workgroup per CU,
// which a row-split launch never exceeds (at most one workgroup per CU).  X ping-pongs between two
// buffers: step g = sweep (nbe - 1) + step reads X (g even) or Xb (g odd)
// and writes the other, so no workgroup can read rows that another
// workgroup of its pair has already rotated (a workgroup may start after
// another has finished; with one buffer, that read raced with the write).
// A pair with nothing to rotate copies its slice across.  A task that has
// converged writes nothing more: its X is in Xb when it ran an odd number of
// sweeps (a sweep is an odd number of steps), and the host copies it back.
// Each slice of W is read and written by its own workgroup only.
template <int NCG, int NCW>
__global__ __launch_bounds__(256, 1) void k_jacobi_fused_r(JacobiArgs a, int sweep, int step) {
  __shared__ PairSmem<8> sm;
  __shared__ PanelLds<NCG, NCW> pl;
  const int t = blockIdx.y;
  int p, q;
  split_block(a.S, p, q);
  int lcol[4], ocol[4];
  thread_cols<8>(a.nbe, step, p, lcol, ocol);
  int xa, xb, wa, wb;
  split_rows16(a.Rh, a.S, q, xa, xb);
  split_rows16(a.Ch, a.S, q, wa, wb);
  const int g = sweep * (a.nbe - 1) + step;
  const double2* Xi = ((g & 1) ? a.Xb : a.X) + (size_t)t * a.xslot;
  double2* Xo = ((g & 1) ? a.X : a.Xb) + (size_t)t * a.xslot;
  double2* W = a.W + (size_t)t * a.wslot;
  // a row-split launch holds few tasks: every chunk loads at entry, for the
  // host's bounds, and is masked once the task's R and C have arrived
  Panel8<NCG> px;
  Panel8<NCW> pw;
  load_panel8<NCG, 0, NCG>(Xi, (size_t)a.ldx, 0, a.Rh, a.ldw, lcol, px);
  load_panel8<NCW, 0, NCW>(W, (size_t)a.ldw, wa, wb, a.ldw, lcol, pw);
  issue_fence();
  const int* inf = a.info + (size_t)t * kInfoLen;
  const int R = inf[kInfoRJ], C = inf[kInfoC];
  if (converged(a, t, sweep)) return;
  const int xe = min(xb, R), we = min(wb, C);
  // (the loads took the host's bounds: nothing to mask when the task fills
  // them and the wave's columns are the task's; uniform over the wave)
  const bool cols_in = lcol[0] < C && lcol[1] < C && lcol[2] < C && lcol[3] < C;
  if (!(cols_in && R >= a.Rh)) mask_panel8_all<NCG>(0, R, C, lcol, px);
  if (!(cols_in && we >= wb)) mask_panel8_all<NCW>(wa, we, C, lcol, pw);
  if (threadIdx.x == 0) sm.rr = 0;  // the solve's round word, ordered by the barrier below
  put_panels8<NCG>(px, pl.v);
  put_panels8<NCW>(pw, pl.v + NCG);
  __syncthreads();
  panel_gram8_lds<NCG>(R, pl.v, sm);  // ends with a barrier after its last use of sm.buf
  int wsel = 0;
  if (panel_solve<8, true, true>(q == 0 ? a.offs + (size_t)sweep * a.nt + t : nullptr,
                                 a.tol_rot, a.tiny2, sm, wsel, inner_step(a, step))) {
    double2 wk[4];
    double wks[4];
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      wk[kk] = sm.Ws(wsel, 4 * kk + (lane >> 4), lane & 15);
      wks[kk] = wk[kk].x + wk[kk].y;
    }
    apply8_tiles_lds<NCG>(Xo, (size_t)a.ldx, xa, xe, W, (size_t)a.ldw, wa, we, C, ocol[0], pl.v,
                          wk, wks);
  } else {
    panel_copy8r<NCG>(Xo, (size_t)a.ldx, xa, xe, C, lcol, px);
  }
}

// ------------------------------------------------------------- launch
// 64-row chunks of the widest slice when n rows split into S slices.
static int slice_chunks(int n, int S) { return (((n + S - 1) / S) + 63) / 64; }

// Block width 8 takes the register-resident kernels when X's panel rows fit
// (one workgroup per pair: up to 256 rows; row split: up to 256 rows per
// slice of X and of W), the streaming ones otherwise.  The gate batches'
// 512-row panels stream: their steps are HBM-bound, and the streaming
// kernel's four workgroups per CU measured as fast as the register-resident
// one's two.
void launch_jacobi_step(const JacobiArgs& a, int npairs, int bw, int sweep, int step, int gy,
                        hipStream_t s) {
  const dim3 grid(npairs, gy), blk(256);
  if (bw == 16)
    hipLaunchKernelGGL(k_jacobi_step<16>, grid, blk, 0, s, a, sweep, step);
  else if (a.Rh <= 128)
    hipLaunchKernelGGL(k_jacobi_step_r<2>, grid, blk, 0, s, a, sweep, step);
  else if (a.Rh <= 256)
    hipLaunchKernelGGL(k_jacobi_step_r<4>, grid, blk, 0, s, a, sweep, step);
  else
    hipLaunchKernelGGL(k_jacobi_step<8>, grid, blk, 0, s, a, sweep, step);
}
void launch_active(const JacobiArgs& a, int sweep, int* act, hipStream_t s) {
  hipLaunchKernelGGL(k_active, dim3(1), dim3(1024), 0, s, a, sweep, act);
}
void launch_jacobi_gram(const JacobiArgs& a, int npairs, int bw, int sweep, int step,
                        hipStream_t s) {
  const dim3 grid(npairs * a.S, a.nt), blk(256);
  const int nc = slice_chunks(a.Rh, a.S);
  if (bw == 16)
    hipLaunchKernelGGL(k_jacobi_gram<16>, grid, blk, 0, s, a, sweep, step);
  else if (nc <= 2)
    hipLaunchKernelGGL(k_jacobi_gram_r<2>, grid, blk, 0, s, a, sweep, step);
  else if (nc <= 4)
    hipLaunchKernelGGL(k_jacobi_gram_r<4>, grid, blk, 0, s, a, sweep, step);
  else
    hipLaunchKernelGGL(k_jacobi_gram<8>, grid, blk, 0, s, a, sweep, step);
}
// Rows of the widest slice of the one-launch kernel (multiples of 16).
static int slice_rows16(int n, int S) { return (((n + S - 1) / S) + 15) / 16 * 16; }

bool jacobi_fused_fits(const JacobiArgs& a, int bw) {
  // (panels of up to 256 rows: a 512-row one spills from the registers;
  // slices of W of up to 128 rows, the kernel's two chunks)
  return bw == 8 && a.S > 1 && a.Rh <= 256 && slice_rows16(a.Ch, a.S) <= 128;
}
void launch_jacobi_fused(const JacobiArgs& a, int npairs, int sweep, int step, hipStream_t s) {
  const dim3 grid(npairs * a.S, a.nt), blk(256);
  if (slice_rows16(a.Ch, a.S) <= 64)
    hipLaunchKernelGGL((k_jacobi_fused_r<4, 1>), grid, blk, 0, s, a, sweep, step);
  else
    hipLaunchKernelGGL((k_jacobi_fused_r<4, 2>), grid, blk, 0, s, a, sweep, step);
}
void launch_jacobi_apply(const JacobiArgs& a, int npairs, int bw, int sweep, int step,
                         hipStream_t s) {
  const dim3 grid(npairs * a.S, a.nt), blk(256);
  const int nc = std::max(slice_chunks(a.Rh, a.S), slice_chunks(a.Ch, a.S));
  if (bw == 16)
    hipLaunchKernelGGL(k_jacobi_apply<16>, grid, blk, 0, s, a, sweep, step);
  else if (nc <= 2)
    hipLaunchKernelGGL(k_jacobi_apply_r<2>, grid, blk, 0, s, a, sweep, step);
  else if (nc <= 4)
    hipLaunchKernelGGL(k_jacobi_apply_r<4>, grid, blk, 0, s, a, sweep, step);
  else
    hipLaunchKernelGGL(k_jacobi_apply<8>, grid, blk, 0, s, a, sweep, step);
}

}  // namespace tnmps
