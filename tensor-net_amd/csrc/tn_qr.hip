// Householder QR preconditioning of the Jacobi SVD for gfx950 (the SVD stands
// in for LAPACK zgesvd, tensor-net/lib/mps.rs:299).
//
// For X (R x C, R >= C, column-major) the library can decompose, instead of
// X itself, the C x C lower-triangular Y = R^H of a QR factorisation of X
// with its columns sorted by decreasing norm:
//   X P = Q R,   Y = R^H,   Y W = U~ (one-sided Jacobi on Y, unchanged),
// so  X = (Q W) Sigma (P U~ Sigma^-1)^H  with Sigma the column norms of U~.
// The SVD's consumers expect X W = U Sigma with W the right singular
// vectors, so the last kernel writes
//   X <- Q (W Sigma)    and    W <- P U~ Sigma^-1
// (a zero column of U~ gives a zero column of W; the rank rule drops it).
// This is the QR preconditioner of Drmac and Veselic (one-sided Jacobi on
// R^H of a column-sorted QR): the rows of R are graded like the sorted
// columns, and Jacobi on such input needs fewer sweeps.  For factored
// updates (C = K < R) the Jacobi problem also shrinks from R to C rows.
//
// The QR is blocked Householder in the LAPACK convention (zgeqrf: H_i =
// I - tau_i v_i v_i^H, Q = H_1 ... H_C; panels of kQrNb columns, Q_panel =
// I - V T V^H with T upper triangular as zlarft builds it).  One launch per
// panel (k_qr_step): block 0 of every task updates the next panel's columns
// with the current panel and factors them at once, in registers (reflector
// c: its column's wave forms it, then every wave applies it to its later
// columns; one workgroup barrier per reflector), and builds T; the other
// blocks update the trailing columns, 16 per workgroup, on the FP64 matrix
// cores: A <- A - V (T^H (V^H A)).  So the next panel's factorisation runs
// beside the rest of the trailing update.
// The first panel and trailing update read X through the sort permutation
// and write the factored matrix E (the V below the diagonal, R on and above
// it) into their own buffer, so no separate gather pass runs.
//
// Lane layout of v_mfma_f64_16x16x4f64 (checked on the device): lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16]; D[i][j] is element i / 4 of lane
// j + 16 (i % 4), i.e. element e of lane l is D[l / 16 + 4 e][l % 16].
#include "tn_device.h"
#include "tn_devmath.h"
#include "tn_kernels.h"

namespace tnmps {

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
__device__ __forceinline__ v4d mfma16(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Orders one wave's LDS accesses (a wave's LDS operations are performed in
// issue order; the barrier keeps the compiler from reordering them).
__device__ __forceinline__ void wave_lds_sync() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}

struct TaskQr {
  int R, C;
  double2* X;
  double2* W;
  double2* E;
  double2* T;
  int* perm0;
  double* cn;
};

// Workspace slot of launch task i (the slot map, or i itself).
__device__ __forceinline__ int qr_slot(const QrArgs& a, int i) { return a.map ? a.map[i] : i; }

__device__ __forceinline__ TaskQr task_qr(const QrArgs& a, int i) {
  const int t = qr_slot(a, i);
  const int* inf = a.info + (size_t)t * kInfoLen;
  TaskQr q;
  q.R = inf[kInfoR];
  q.C = inf[kInfoC];
  q.X = a.X + (size_t)t * a.xslot;
  q.W = a.W + (size_t)t * a.wslot;
  q.E = a.E + (size_t)t * a.xslot;
  q.T = a.T + (size_t)t * a.tslot;
  q.perm0 = a.perm0 + (size_t)t * a.ldw;
  q.cn = a.cn + (size_t)t * a.ldw;
  return q;
}

// Column c of the matrix a panel or trailing update reads: panel 0's pass
// reads X through the sort order, later ones the factored matrix E.
__device__ __forceinline__ const double2* src_col(const QrArgs& a, const TaskQr& q, int jp,
                                                  int c) {
  return jp == 0 ? q.X + (size_t)q.perm0[c] * a.ldx : q.E + (size_t)c * a.ldx;
}

}  // namespace

// ------------------------------------------------------------ column norms
// Squared column norms of X (16 columns per workgroup, 4 per wave).
__global__ __launch_bounds__(256) void k_qr_norms(QrArgs a) {
  const TaskQr q = task_qr(a, blockIdx.y);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int j = blockIdx.x * 16 + wid * 4 + k;
    if (j >= q.C) break;
    const double2* col = q.X + (size_t)j * a.ldx;
    double s = 0.0;
    for (int r = lane; r < q.R; r += 64) s += cabs2(col[r]);
    s = wave_sum_dpp(s);
    if (lane == 0) q.cn[j] = s;
  }
}

// ------------------------------------------------------------ sort order
// perm0: columns by decreasing norm, ties by increasing index (one
// workgroup per task).
__global__ __launch_bounds__(1024) void k_qr_sort(QrArgs a) {
  const TaskQr q = task_qr(a, blockIdx.x);
  __shared__ double key[kMaxCols];
  __shared__ int idx[kMaxCols];
  int P2 = 1;
  while (P2 < q.C) P2 <<= 1;
  for (int i = threadIdx.x; i < P2; i += blockDim.x) {
    key[i] = (i < q.C) ? q.cn[i] : -1.0;
    idx[i] = i;
  }
  __syncthreads();
  bitonic_desc(key, idx, P2);
  for (int i = threadIdx.x; i < q.C; i += blockDim.x) q.perm0[i] = idx[i];
}

// ------------------------------------------------------- null columns
// Columns (below the diagonal) with norm at most kQrNullCol get no
// reflector: far below f64::EPSILON times the SVD input's norm (about 1),
// they cannot change a kept singular value, and beta |alpha - beta|^2
// stays a normal number above it.
constexpr double kQrNullCol = 1e-90;

// ------------------------------------------------- one panel on registers
// The trailing update and the final Q application run one panel's
// reflectors, Z <- (I - V op(T) V^H) Z, on a 16-column tile Z held in
// registers by a workgroup of kQW waves, in the matrix-core output layout:
// wave w holds row tiles t = w, w + kQW, ... (16 rows each; R <=
// kQrMaxRows), and element e of lane l of tile t is Z[16 t + l / 16 + 4 e]
// [l % 16].  That layout is also the operand of V^H Z computed as
// (Z^H V)^H, so only V changes layout: it is loaded with lanes along rows
// (the A operand of V M) and transposed through a per-wave LDS tile (the B
// operand of Z^H V).  The kQW partial Z^H V are summed in a fixed order.
namespace {
constexpr int kQW = 8;
constexpr int kQT = kQrMaxRows / 16 / kQW;  // row tiles per wave

struct PanelSmem {
  double2 part[kQW][kQrNb][kQrNb + 1];
  double2 Ms[kQrNb][kQrNb + 1];
  double2 Ts[kQrNb][kQrNb + 1];
  double2 vt[kQW][kQrNb][kQrNb + 1];
};

// V[r][i] of a panel for local row r (global g0 + r) and panel column i:
// 1 on the diagonal, 0 above it, E below it; 0 past the panel's ncol
// reflectors and past its nr rows.
// (zero as well where `use` is false)
__device__ __forceinline__ double2 vload(const double2* E, size_t ldx, int g0, int nr, int ncol,
                                         int r, int i, bool use = true) {
  const bool below = use && r < nr && r > i && i < ncol;
  const double2 v = ld_or_zero(E + (size_t)(g0 + i) * ldx + g0 + r, below, E);
  return (use && r == i && i < ncol && r < nr) ? make_double2(1.0, 0.0) : v;
}

// Panel jp's V for the wave's row tiles, lanes along rows: lane r + 16 k of
// slot (s, kk) holds V[16 t + r][4 kk + k] (zero for tiles above the panel).
__device__ __forceinline__ void load_vrows(const double2* E, size_t ldx, int R, int C, int jp,
                                           double2 (&v)[kQT][4]) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int g0 = kQrNb * jp, nr = R - g0, ncol = min(kQrNb, C - g0);
#pragma unroll
  for (int s = 0; s < kQT; ++s) {
    const int t = wid + kQW * s;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
      v[s][kk] = vload(E, ldx, g0, nr, ncol, 16 * t + (lane & 15) - g0, 4 * kk + (lane >> 4),
                       t >= jp);
  }
}

// Z <- (I - V op(T) V^H) Z with op(T) = T^H (ADJ: Q_panel^H, the QR's
// trailing update) or T (Q_panel, the final application of Q).  Tp: the
// panel's T, row-major.  Begins with a workgroup barrier.
template <bool ADJ>
__device__ void panel_update(const double2 (&vrow)[kQT][4], const double2* Tp, int jp, int R,
                             v4d (&zr)[kQT], v4d (&zi)[kQT], PanelSmem& sm) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  __syncthreads();  // the previous use of sm is done
  if (tid < kQrNb * kQrNb) sm.Ts[tid >> 4][tid & 15] = Tp[tid];
  // H = Z^H V over the wave's row tiles at or below the panel, with three
  // real products instead of four: P1 = Zr^T Vr, P2 = Zi^T Vi,
  // P3 = (Zr - Zi)^T (Vr + Vi); Re H = P1 + P2, Im H = P3 - P1 + P2
  v4d p1 = {0.0, 0.0, 0.0, 0.0}, p2 = {0.0, 0.0, 0.0, 0.0}, p3 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < kQT; ++s) {
    const int t = wid + kQW * s;
    if (t >= jp && 16 * t < R) {
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) sm.vt[wid][lane & 15][4 * kk + (lane >> 4)] = vrow[s][kk];
      wave_lds_sync();
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        // B operand: lane i + 16 k holds V[16 t + 4 e + k][i]; A operand
        // conj(Z)^T: lane j + 16 k holds Z[16 t + 4 e + k][j] (element e)
        const double2 v = sm.vt[wid][4 * e + (lane >> 4)][lane & 15];
        p1 = mfma16(zr[s][e], v.x, p1);
        p2 = mfma16(zi[s][e], v.y, p2);
        p3 = mfma16(zr[s][e] - zi[s][e], v.x + v.y, p3);
      }
      wave_lds_sync();
    }
  }
  const v4d hre = p1 + p2, him = p3 - p1 + p2;
#pragma unroll
  for (int e = 0; e < 4; ++e)
    sm.part[wid][(lane >> 4) + 4 * e][lane & 15] = make_double2(hre[e], him[e]);
  __syncthreads();
  if (tid < kQrNb * kQrNb) {
    // H summed over the waves in a fixed order, into part[0]
    const int j = tid >> 4, l = tid & 15;
    double2 h = sm.part[0][j][l];
#pragma unroll
    for (int w = 1; w < kQW; ++w) h = cadd(h, sm.part[w][j][l]);
    sm.part[0][j][l] = h;
  }
  __syncthreads();
  if (tid < kQrNb * kQrNb) {
    // (V^H Z)[l][j] = conj(H[j][l]);  M = op(T) V^H Z
    const int i = tid >> 4, j = tid & 15;
    double2 acc = make_double2(0.0, 0.0);
    for (int l = ADJ ? 0 : i; l < (ADJ ? i + 1 : kQrNb); ++l) {
      const double2 t = ADJ ? cconj(sm.Ts[l][i]) : sm.Ts[i][l];
      acc = cadd(acc, cmul(t, cconj(sm.part[0][j][l])));
    }
    sm.Ms[i][j] = acc;
  }
  __syncthreads();
  double2 mk[4];
  double mks[4];
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    mk[kk] = sm.Ms[4 * kk + (lane >> 4)][lane & 15];
    mks[kk] = mk[kk].x + mk[kk].y;
  }
  // Z -= V M:  A operand V (lane r + 16 k: V[r][4 kk + k]), B operand M;
  // three real products per tile: Q1 = Vr Mr, Q2 = Vi Mi,
  // Q3 = (Vr + Vi)(Mr + Mi); Re VM = Q1 - Q2, Im VM = Q3 - Q1 - Q2
#pragma unroll
  for (int s = 0; s < kQT; ++s) {
    const int t = wid + kQW * s;
    if (t >= jp && 16 * t < R) {
      v4d q1 = {0.0, 0.0, 0.0, 0.0}, q2 = {0.0, 0.0, 0.0, 0.0}, q3 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const double2 v = vrow[s][kk];
        q1 = mfma16(v.x, mk[kk].x, q1);
        q2 = mfma16(v.y, mk[kk].y, q2);
        q3 = mfma16(v.x + v.y, mks[kk], q3);
      }
      zr[s] -= q1 - q2;
      zi[s] -= q3 - q1 - q2;
    }
  }
}
}  // namespace

// ------------------------------------------- QR step with look-ahead
// One launch per panel: block 0 of every task updates the next panel's
// columns with the current panel and factors them at once (so the next
// panel's factorisation runs beside the rest of the trailing update), the
// other blocks update the trailing columns beyond.  Launch jp = -1 factors
// panel 0 only.
namespace {
struct FactorSmem {
  // reflector c's vector, in buffer c & 1 (a buffer is rewritten two
  // reflectors later, after every wave has passed the barrier between)
  double2 vbuf[2][kQrMaxRows];
  double2 taus[kQrNb];
  double2 g[kQrNb][kQrNb + 1];
};

constexpr int kRPL = kQrMaxRows / 64;  // rows per lane

// Reflector c from the owning wave's column xc (zlarfg): alpha = x[c],
// xnorm = |x[c+1:]|, beta = -sign(Re alpha) sqrt(|alpha|^2 + xnorm^2),
// tau = (beta - alpha) / beta, v = x[c+1:] / (alpha - beta), x[c] = beta
// (tau = 0 and nothing changes when xnorm = 0 and Im alpha = 0, or when the
// column is at most kQrNullCol); one division gives 1 / beta and
// 1 / |alpha - beta|^2.  Posted to fs.vbuf[c & 1] and fs.taus[c].
__device__ __forceinline__ void reflect(double2 (&xc)[kRPL], double s, int c, int nr,
                                        FactorSmem& fs) {
  const int lane = threadIdx.x & 63;
  const double2 alpha = make_double2(read_lane(xc[0].x, c), read_lane(xc[0].y, c));
  double2 tau = make_double2(0.0, 0.0);
  const double mag = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + s);
  if ((s > 0.0 || alpha.y != 0.0) && mag > kQrNullCol) {
    const double beta = alpha.x >= 0.0 ? -mag : mag;
    const double2 d = make_double2(alpha.x - beta, alpha.y);
    const double d2 = cabs2(d);
    const double rr = 1.0 / (beta * d2);
    const double ibeta = d2 * rr, id2 = beta * rr;
    tau = make_double2((beta - alpha.x) * ibeta, -alpha.y * ibeta);
    const double2 sc = make_double2(d.x * id2, -d.y * id2);  // 1 / d
#pragma unroll
    for (int m = 0; m < kRPL; ++m)
      if (lane + 64 * m > c) xc[m] = cmul(xc[m], sc);
    if (lane == c) xc[0] = make_double2(beta, 0.0);
  }
#pragma unroll
  for (int m = 0; m < kRPL; ++m)
    if (lane + 64 * m < nr) fs.vbuf[c & 1][lane + 64 * m] = xc[m];
  if (lane == 0) fs.taus[c] = tau;
}

// Panel jp of task q factored by the 512-thread workgroup (R - 16 jp <=
// kQrMaxRows rows): wave w holds
// panel columns w and w + 8 in registers (lane l: rows l + 64 m).  Per
// reflector c the owning wave forms it; after one barrier every wave forms
// w = v_c^H x for both its columns (one pass, the two sums interleaved on
// the DPP network), updates the later columns (x -= conj(tau) w v) and,
// for columns before c, records (V^H V)[col][c] = conj(w) for T.  The
// column next in line forms its norm during its update.  T (zlarft,
// forward, columnwise): T[i][i] = tau_i,
//   T[j][i] = -tau_i sum_{l=j}^{i-1} T[j][l] (V^H V)[l][i]   (j < i),
// one row per lane of wave 0 (a row depends only on its own earlier entries).
__device__ void factor_panel8(const QrArgs& a, const TaskQr& q, int jp, FactorSmem& fs) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int g0 = kQrNb * jp;
  const int nr = q.R - g0;
  const int ncol = min(kQrNb, q.C - g0);
  double2 x[2][kRPL];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int col = wid + 8 * j;
    const double2* src = col < ncol ? src_col(a, q, jp, g0 + col) + g0 : nullptr;
#pragma unroll
    for (int m = 0; m < kRPL; ++m) {
      const int r = lane + 64 * m;
      x[j][m] = ld_or_zero(src + r, src && r < nr, q.E);
    }
  }
  double s_own[2] = {0.0, 0.0};
  if (wid == 0) {
    double sl = 0.0;
#pragma unroll
    for (int m = 0; m < kRPL; ++m)
      if (lane + 64 * m > 0) sl += cabs2(x[0][m]);
    s_own[0] = wave_sum_dpp(sl);
  }
  for (int c = 0; c < ncol; ++c) {
    if (wid == (c & 7)) {
      if (c < 8)
        reflect(x[0], s_own[0], c, nr, fs);
      else
        reflect(x[1], s_own[1], c, nr, fs);
    }
    __syncthreads();
    const double2 tau = fs.taus[c];
    double2 v[kRPL];
    double2 w0 = make_double2(0.0, 0.0), w1 = make_double2(0.0, 0.0);
#pragma unroll
    for (int m = 0; m < kRPL; ++m) {
      const int r = lane + 64 * m;
      v[m] = (r > c && r < nr) ? fs.vbuf[c & 1][r] : make_double2(r == c ? 1.0 : 0.0, 0.0);
      w0 = cadd(w0, cmulc(v[m], x[0][m]));
      w1 = cadd(w1, cmulc(v[m], x[1][m]));
    }
    w0 = wave_csum_dpp(w0);
    w1 = wave_csum_dpp(w1);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = wid + 8 * j;
      const double2 w = j == 0 ? w0 : w1;
      if (col > c && col < ncol) {
        const double2 f = cmulc(tau, w);  // conj(tau) w
        double sl = 0.0;
#pragma unroll
        for (int m = 0; m < kRPL; ++m) {
          const double2 fv = cmul(f, v[m]);
          x[j][m] = make_double2(x[j][m].x - fv.x, x[j][m].y - fv.y);
          if (lane + 64 * m > c + 1) sl += cabs2(x[j][m]);
        }
        if (col == c + 1) s_own[j] = wave_sum_dpp(sl);
      } else if (col < c && lane == 0) {
        fs.g[col][c] = cconj(w);  // (V^H V)[col][c] over rows >= c
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int col = wid + 8 * j;
    if (col < ncol) {
      double2* dst = q.E + (size_t)(g0 + col) * a.ldx + g0;
#pragma unroll
      for (int m = 0; m < kRPL; ++m) {
        const int r = lane + 64 * m;
        if (r < nr) dst[r] = x[j][m];
      }
    }
  }
  __syncthreads();
  // T in two 8 x 8 diagonal blocks and the block above the diagonal,
  // T = [[T1, T12], [0, T2]], T12 = -T1 (V^H V)[0:8][8:16] T2 (the product
  // of two compact-WY factors), on wave 0 through LDS (the reflector
  // buffers, free once every wave has passed the barrier above): lane j < 16
  // forms row j of its diagonal block by the recurrence, then 64 lanes form
  // P = T1 G12 and T12 = -P T2 with at most 8 terms each.  Columns at or
  // past ncol are zero, and no product reads a (V^H V) entry of them.
  if (wid == 0) {
    double2(*tt)[kQrNb + 1] = reinterpret_cast<double2(*)[kQrNb + 1]>(&fs.vbuf[0][0]);
    double2(*pb)[9] = reinterpret_cast<double2(*)[9]>(&fs.vbuf[1][0]);
    if (lane < kQrNb) {
      const int j = lane, o = j & 8;
      double2 trow[8];
#pragma unroll
      for (int ii = 0; ii < 8; ++ii) {
        const int i = o + ii;
        double2 v = make_double2(0.0, 0.0);
        if (i < ncol) {
          if (i == j) {
            v = fs.taus[i];
          } else if (i > j) {
            double2 sm = make_double2(0.0, 0.0);
#pragma unroll
            for (int ll = 0; ll < 8; ++ll)
              if (o + ll >= j && ll < ii) sm = cadd(sm, cmul(trow[ll], fs.g[o + ll][i]));
            const double2 ts = cmul(fs.taus[i], sm);
            v = make_double2(-ts.x, -ts.y);
          }
        }
        trow[ii] = v;
      }
#pragma unroll
      for (int ii = 0; ii < 8; ++ii) {
        tt[j][o + ii] = trow[ii];
        tt[j][(o ^ 8) + ii] = make_double2(0.0, 0.0);
      }
    }
    wave_lds_sync();
    const int r = lane >> 3, m = lane & 7;
    {
      double2 acc = make_double2(0.0, 0.0);
      if (8 + m < ncol) {
#pragma unroll
        for (int l = 0; l < 8; ++l)
          if (l >= r) acc = cadd(acc, cmul(tt[r][l], fs.g[l][8 + m]));
      }
      pb[r][m] = acc;
    }
    wave_lds_sync();
    {
      double2 acc = make_double2(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k <= m) acc = cadd(acc, cmul(pb[r][k], tt[8 + k][8 + m]));
      tt[r][8 + m] = make_double2(-acc.x, -acc.y);
    }
    wave_lds_sync();
    double2* T = q.T + (size_t)jp * kQrNb * kQrNb;
#pragma unroll
    for (int e4 = 0; e4 < 4; ++e4) {
      const int e = lane + 64 * e4;
      T[e] = tt[e >> 4][e & 15];
    }
  }
}

// Columns c0 .. c0 + 15 of task q, rows 16 jp .. R-1: A <- A - V (T^H (V^H A))
// with panel jp, on registers (panel_update), written to E.
__device__ void trail_tile(const QrArgs& a, const TaskQr& q, int jp, int c0, PanelSmem& sm) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int col = c0 + (lane & 15);
  const bool colok = col < q.C;
  const double2* src = colok ? src_col(a, q, jp, col) : nullptr;
  v4d zr[kQT], zi[kQT];
#pragma unroll
  for (int s = 0; s < kQT; ++s) {
    const int t = wid + kQW * s;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = 16 * t + (lane >> 4) + 4 * e;
      const double2 x = ld_or_zero(src + row, colok && t >= jp && row < q.R, q.E);
      zr[s][e] = x.x;
      zi[s][e] = x.y;
    }
  }
  double2 vrow[kQT][4];
  load_vrows(q.E, a.ldx, q.R, q.C, jp, vrow);
  panel_update<true>(vrow, q.T + (size_t)jp * kQrNb * kQrNb, jp, q.R, zr, zi, sm);
  if (colok) {
    double2* dst = q.E + (size_t)col * a.ldx;
#pragma unroll
    for (int s = 0; s < kQT; ++s) {
      const int t = wid + kQW * s;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = 16 * t + (lane >> 4) + 4 * e;
        if (t >= jp && row < q.R) dst[row] = make_double2(zr[s][e], zi[s][e]);
      }
    }
  }
}
}  // namespace

__global__ __launch_bounds__(512) void k_qr_step(QrArgs a, int jp) {
  __shared__ PanelSmem sm;
  __shared__ FactorSmem fs;
  const TaskQr q = task_qr(a, blockIdx.y);
  if (q.R > kQrMaxRows) return;  // (the host takes this path for R <= kQrMaxRows)
  const int np = jp + 1;  // the panel block 0 factors
  if (blockIdx.x == 0) {
    if (kQrNb * np >= q.C) return;
    if (jp >= 0) {
      trail_tile(a, q, jp, kQrNb * np, sm);
      __syncthreads();  // the updated columns, in E, are read back below
    }
    factor_panel8(a, q, np, fs);
    return;
  }
  const int c0 = kQrNb * (np + blockIdx.x);
  if (c0 >= q.C) return;
  trail_tile(a, q, jp, c0, sm);
}

// ------------------------------------------------------------ Y = R^H
// Y (C x C, column-major in X's slot, ld ldx) = R^H: Y[j][i] = conj(R[i][j])
// = conj(E[j][i] stored at column j, row i) for i <= j, zero above the
// diagonal.  32 x 32 tiles through LDS; the Jacobi problem's row count
// becomes C.
__global__ __launch_bounds__(256) void k_qr_formy(QrArgs a, int ntile) {
  __shared__ double2 tile[32][33];
  const TaskQr q = task_qr(a, blockIdx.y);
  const int bi = blockIdx.x % ntile, bj = blockIdx.x / ntile;
  if (blockIdx.x == 0 && threadIdx.x == 0)
    a.info[(size_t)qr_slot(a, blockIdx.y) * kInfoLen + kInfoRJ] = q.C;
  const int i0 = 32 * bi, j0 = 32 * bj;  // R rows i, columns j
  if (i0 >= q.C || j0 >= q.C) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int jj = ty; jj < 32; jj += 8) {
    const int i = i0 + tx, j = j0 + jj;
    tile[jj][tx] = (i < q.C && j < q.C && i <= j) ? q.E[(size_t)j * a.ldx + i]
                                                  : make_double2(0.0, 0.0);
  }
  __syncthreads();
  // Y column i (= R row i), rows j
  for (int ii = ty; ii < 32; ii += 8) {
    const int i = i0 + ii, j = j0 + tx;
    if (i < q.C && j < q.C) q.X[(size_t)i * a.ldx + j] = cconj(tile[tx][ii]);
  }
}

// ------------------------------------------------------------ final
// After the Jacobi sweeps on Y (X's slot holds U~ = Y W): columns j0 .. j0+15
// of every task, one workgroup of kQW waves.
//   sigma_j = |U~[:, j]|;  W[:, j] <- P U~[:, j] / sigma_j  (row perm0[i] <- row i);
//   Z = [W[:, j] sigma_j ; 0] (R x 16)  <-  Q Z, applying the panels last to
//   first (panel_update);  X[:, j] <- Z.
namespace {
struct FinSmem {
  PanelSmem p;
  double s2[kQW][kQrNb];
  double sig[kQrNb];
};
}  // namespace

__global__ __launch_bounds__(512) void k_qr_final(QrArgs a) {
  __shared__ FinSmem sm;
  const TaskQr q = task_qr(a, blockIdx.y);
  const int j0 = kQrNb * blockIdx.x;
  if (j0 >= q.C || q.R > kQrMaxRows) return;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int col = j0 + (lane & 15);
  const bool colok = col < q.C;
  // U~ column tile: squared norms, kept for the new W
  double2 u[kQT][4];
  double s2 = 0.0;
#pragma unroll
  for (int s = 0; s < kQT; ++s) {
    const int t = wid + kQW * s;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = 16 * t + (lane >> 4) + 4 * e;
      u[s][e] = ld_or_zero(q.X + (size_t)col * a.ldx + row, colok && row < q.C, q.X);
      s2 += cabs2(u[s][e]);
    }
  }
  s2 += __shfl_xor(s2, 16, 64);
  s2 += __shfl_xor(s2, 32, 64);
  // W column tile, the start of Z
  v4d zr[kQT], zi[kQT];
#pragma unroll
  for (int s = 0; s < kQT; ++s) {
    const int t = wid + kQW * s;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = 16 * t + (lane >> 4) + 4 * e;
      const double2 w = ld_or_zero(q.W + (size_t)col * a.ldw + row, colok && row < q.C, q.W);
      zr[s][e] = w.x;
      zi[s][e] = w.y;
    }
  }
  if (lane < 16) sm.s2[wid][lane] = s2;
  __syncthreads();
  if (tid < kQrNb) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kQW; ++w) v += sm.s2[w][tid];
    sm.sig[tid] = sqrt(v);
  }
  __syncthreads();
  const double sg = sm.sig[lane & 15];
  // (columns of U~ that small are dropped by the rank rule; their W stays 0)
  const double isg = sg > 1e-150 ? 1.0 / sg : 0.0;
#pragma unroll
  for (int s = 0; s < kQT; ++s) {
    zr[s] *= sg;
    zi[s] *= sg;
  }
  // the scatter targets, loaded together (index 0 where unused)
  int prow[kQT][4];
#pragma unroll
  for (int s = 0; s < kQT; ++s) {
    const int t = wid + kQW * s;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = 16 * t + (lane >> 4) + 4 * e;
      prow[s][e] = q.perm0[(colok && row < q.C) ? row : 0];
    }
  }
  // every read of this tile of W (in Z) precedes the writes below
  __syncthreads();
#pragma unroll
  for (int s = 0; s < kQT; ++s) {
    const int t = wid + kQW * s;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = 16 * t + (lane >> 4) + 4 * e;
      if (colok && row < q.C) q.W[(size_t)col * a.ldw + prow[s][e]] = cscale(u[s][e], isg);
    }
  }
  // Z <- Q Z, panels last to first
  const int npan = (q.C + kQrNb - 1) / kQrNb;
  for (int jp = npan - 1; jp >= 0; --jp) {
    double2 vrow[kQT][4];
    load_vrows(q.E, a.ldx, q.R, q.C, jp, vrow);
    panel_update<false>(vrow, q.T + (size_t)jp * kQrNb * kQrNb, jp, q.R, zr, zi, sm.p);
  }
  // X[:, j] <- Z (rows < R)
  if (colok) {
#pragma unroll
    for (int s = 0; s < kQT; ++s) {
      const int t = wid + kQW * s;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = 16 * t + (lane >> 4) + 4 * e;
        if (row < q.R) q.X[(size_t)col * a.ldx + row] = make_double2(zr[s][e], zi[s][e]);
      }
    }
  }
  // The column norms of X for the singular values (k_finish), while this
  // block's columns are in its cache, with k_finish's arithmetic: wave w
  // sums rows lane, lane + 64, ... of columns j0 + w and j0 + w + 8 in that
  // order, then the same lane tree, so the norms are bitwise those k_finish
  // would form (it skips its pass for the task).  sraw shares cn's slots.
  __syncthreads();  // every column of X above written
  for (int h = 0; h < 2; ++h) {
    const int j = j0 + wid + (kQrNb / 2) * h;
    if (j < q.C) {
      double s = 0.0;
      for (int r = lane; r < q.R; r += 64) s += cabs2(q.X[(size_t)j * a.ldx + r]);
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
      if (lane == 0) q.cn[j] = sqrt(s);
    }
  }
  if (blockIdx.x == 0 && tid == 0)
    a.info[(size_t)qr_slot(a, blockIdx.y) * kInfoLen + kInfoNorms] = 1;
}

// ------------------------------------------------------------- launch
void launch_qr(const QrArgs& a, int maxR, int maxC, hipStream_t s) {
  const int ctiles = (maxC + kQrNb - 1) / kQrNb;
  hipLaunchKernelGGL(k_qr_norms, dim3(ctiles, a.nt), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_qr_sort, dim3(a.nt), dim3(1024), 0, s, a);
  // launch jp factors panel jp + 1 (block 0) and updates the columns after
  // it with panel jp
  for (int jp = -1; jp + 1 < ctiles; ++jp) {
    const int rest = jp < 0 ? 0 : ctiles - jp - 2;
    hipLaunchKernelGGL(k_qr_step, dim3(1 + rest, a.nt), dim3(512), 0, s, a, jp);
  }
  const int yt = (maxC + 31) / 32;
  hipLaunchKernelGGL(k_qr_formy, dim3(yt * yt, a.nt), dim3(256), 0, s, a, yt);
  (void)maxR;
}

void launch_qr_final(const QrArgs& a, int maxC, hipStream_t s) {
  const int ctiles = (maxC + kQrNb - 1) / kQrNb;
  hipLaunchKernelGGL(k_qr_final, dim3(ctiles, a.nt), dim3(512), 0, s, a);
}

}  // namespace tnmps
