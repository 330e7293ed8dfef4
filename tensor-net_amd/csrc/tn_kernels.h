// Host-side launchers for the kernels in tn_kernels.hip and tn_jacobi.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "tn_device.h"

namespace tnmps {

// theta on the matrix cores in tm x tm tiles (tm = 32 or 64); gx workgroups
// per task loop over the task's ceil(chiL / (tm/2)) * ceil(chiR / (tm/2)) tiles.
void launch_theta(const UpdDesc* descs, int nt, int tm, int gx, hipStream_t s);
void launch_winit(const UpdDesc* descs, int nt, int gx, hipStream_t s);
// Jacobi step with block width bw (8 or 16): one workgroup per block pair
// and task (grid rows gy = a.nt), or, with a.act set, per block pair and
// listed task (gy at least the list's length).
void launch_jacobi_step(const JacobiArgs& a, int npairs, int bw, int sweep, int step, int gy,
                        hipStream_t s);
// act[0] = tasks not converged before `sweep`, act[1 ..] their indices in
// ascending order (one workgroup; a.nt <= 8192).
void launch_active(const JacobiArgs& a, int sweep, int* act, hipStream_t s);
// Row-split step: a.S workgroups per block pair (gram, then solve + apply).
void launch_jacobi_gram(const JacobiArgs& a, int npairs, int bw, int sweep, int step,
                        hipStream_t s);
void launch_jacobi_apply(const JacobiArgs& a, int npairs, int bw, int sweep, int step,
                         hipStream_t s);
// Row-split step in one launch (every workgroup of a pair forms the whole
// panel's Gram matrix and writes its slice of the rotated rows into the
// other X buffer: step g = sweep (nbe - 1) + step reads a.X when g is even,
// a.Xb when odd); jacobi_fused_fits says whether the panels fit its
// registers.
bool jacobi_fused_fits(const JacobiArgs& a, int bw);
void launch_jacobi_fused(const JacobiArgs& a, int npairs, int sweep, int step, hipStream_t s);
// QR preconditioner (tn_qr.hip): sort order, blocked Householder QR of the
// sorted X into E, and Y = R^H into X's slot (info[kInfoRJ] = C); after the
// Jacobi sweeps on Y, X <- Q (W Sigma) and W <- P U~ Sigma^-1.  Every task
// has R <= kQrMaxRows; maxC bounds the tasks' C.
void launch_qr(const QrArgs& a, int maxR, int maxC, hipStream_t s);
void launch_qr_final(const QrArgs& a, int maxC, hipStream_t s);
void launch_finish(const UpdDesc* descs, int nt, int sweeps_run, hipStream_t s);
// Factored updates: T = kept singular vectors x isometric factor; grid bounds
// maxX >= max(m, n) and maxRank >= rank over the tasks.
void launch_fmul(const UpdDesc* descs, int nt, int maxX, int maxRank, hipStream_t s);
void launch_split(const UpdDesc* descs, int nt, hipStream_t s);
void launch_renorm(const UpdDesc* descs, int nt, hipStream_t s);
// G workgroups per site (1 .. kMeasParts); parts holds 2 kMeasParts doubles per site.
void launch_measure(const MeasDesc* descs, int nm, int G, double* parts, hipStream_t s);
// One workgroup per entropy read-out descs[i], result to out[descs[i].slot].
void launch_entropy(const EntDesc* descs, int count, double* out, hipStream_t s);
// One site of the state-vector contraction: out (2P x chiR) from acc
// (P x chiL); max_out bounds 2P x chiR for the grid.
void launch_contract(const double2* acc, size_t P, const int* dims, const double2* g,
                     const double* lamL, size_t max_out, double2* out, hipStream_t s);
void launch_u1(double2* g, const int* dims, const double* op8, int gx, hipStream_t s);
// Factorization of a state vector (mps.rs:448-508), one site per call:
// info and offs of the SVD task, W = I, and for orientation B X = q from
// conj(q) (k_fv_init); then, after the Jacobi sweeps and k_finish with the
// rank known, Gamma[k] and the next factor (k_fv_split); the last site from
// the final factor (k_fv_last, m/2 rows of 2 values).
void launch_fv_init(const FvArgs& a, hipStream_t s);
// rho (D x D, row-major) = sum_(l, r) lamR[r]^2 T[l][i][r] conj(T[l][j][r]) over
// T of L x D x R; lamR null means weights 1.
void launch_rho_part(const double2* t, size_t L, size_t D, size_t R, const double* lamR,
                     double2* rho, hipStream_t s);
void launch_fv_split(const FvArgs& a, hipStream_t s);
void launch_fv_last(const double2* q, const double* lam, int udim, double2* g, hipStream_t s);
// Haar 4x4 unitaries from the counter-based generator (tn_rng.h): unitary u
// from counter words ids[3u .. 3u+2], written row-major to out + 16 u.
void launch_haar4(uint64_t seed, const uint64_t* ids, int count, double2* out, hipStream_t s);

}  // namespace tnmps
