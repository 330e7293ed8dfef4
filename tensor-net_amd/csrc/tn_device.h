// Device-side data structures shared by the kernels and the host driver of
// the Schmidt-form (Gamma-Lambda) MPS update path.
//
// Layout in HBM (one MPS, "chain"):
//   gamma : n site slots.  Site k holds Gamma[k] with axes [u_{k-1}, s_k, u_k]
//           (tensor-net/lib/mps.rs:345-350), row-major and packed by its
//           current dimensions (dims[k], 2, dims[k+1]); the slot is sized for
//           the largest dimensions the bond can reach.
//   lam   : n-1 bond slots of doubles; bond k holds lambda[k] (mps.rs:352).
//   dims  : n+1 ints; dims[0] = dims[n] = 1, dims[k+1] = dimension of bond k.
//
// One two-site update (contract -> SVD -> truncate -> split -> renormalise)
// runs as a "task".  theta = A M with A = lamL Gamma[k] (m x K) and
// M = lamM Gamma[k+1] lamR (K x n), m = 2 chiL, K = chiM, n = 2 chiR.
// In the full mode the SVD works on X, the contracted theta or its conjugate
// transpose, whichever has at least as many rows as columns, stored
// column-major with R = max(m, n) rows and C = min(m, n) columns.  In the
// factored modes one factor is known to be an isometry (its site was last
// written from the singular vectors of an update and not touched since), and
// K < min(m, n): the SVD of theta is that of the other factor, so X is
//   mode 1 (lamL Gamma[k] has orthonormal columns):   X = M^H (n x K)
//   mode 2 (Gamma[k+1] lamR has orthonormal rows):    X = A lamM (m x K)
// and the isometric factor is multiplied into the new Gamma afterwards.
// W (C x C, column-major) accumulates the one-sided Jacobi rotations: on
// exit X W has orthogonal columns.
#pragma once
#include <hip/hip_runtime.h>

namespace tnmps {

constexpr int kBW = 8;          // Jacobi block width for batched work (columns)
constexpr int kBWWide = 16;     // Jacobi block width for lone large SVDs
constexpr int kMaxSweeps = 64;  // outer Jacobi sweep limit (slow graded inputs take up to ~45)
constexpr int kNPart = 64;      // blocks of the split/renormalise kernels
constexpr int kMaxCols = 1024;  // largest C (= min(m, n)) the SVD accepts
constexpr int kMaxSplit = 8;    // workgroups per block pair in the row-split step
constexpr int kFusedSlots = 4;  // tasks whose row-split steps can take one launch (second X buffers)
constexpr int kMeasParts = 64;  // most workgroups per measured site
constexpr int kQrNb = 16;       // Householder panel width of the QR preconditioner
constexpr int kQrMaxRows = 512; // most rows R of an SVD the QR preconditioner takes

// info[] slots of a task (device ints)
enum {
  kInfoM = 0, kInfoN, kInfoR, kInfoC, kInfoOrient, kInfoRank, kInfoFlags,
  kInfoSweeps, kInfoChiL, kInfoChiM, kInfoChiR, kInfoMode,
  kInfoRJ,  // rows of the Jacobi problem: R, or C after the QR preconditioner
  kInfoNorms,  // 1 when X's column norms are already in sraw (k_qr_final)
  kInfoLen = 16
};

// failure flag bits (the values of TN_FLAG_* in tnmps.h): in info[kInfoFlags]
// and the handle's and engine's sticky words
constexpr int kFlagNoConv = 1;     // Jacobi stopped at the sweep limit (set by the host)
constexpr int kFlagCapacity = 2;   // rank clamped to the bond's allocation (k_finish)

// One two-site update (mps.rs:846-880 local_refactor, or mps.rs:1332-1387
// apply_unitary2 when has_gate is set).  All pointers are device addresses.
struct UpdDesc {
  double2* gk;           // Gamma[k]
  double2* gk1;          // Gamma[k+1]
  const double* lamL;    // lambda[k-1]; nullptr when k == 0
  double* lamM;          // lambda[k], overwritten by the new Schmidt values
  const double* lamR;    // lambda[k+1]; nullptr when k == n-2
  int* dims;             // &dims[k]: chiL = dims[0], chiM = dims[1], chiR = dims[2]
  double2* X;            // SVD working matrix, ldx rows
  double2* W;            // accumulated right rotation, ldw rows
  double* sraw;          // column norms of X after convergence (C)
  double* sval;          // new lambda values, sorted (C)
  double* fac;           // scale of the right factor rows (C)
  int* perm;             // sorted column order (C)
  double* part;          // per-block partial norms (2 * kNPart)
  double* offs;          // per-sweep max off-diagonal ratio, stride nt
  int* info;             // kInfoLen ints
  double2* T;            // factored modes: product of the kept singular vectors
                         // with the isometric factor (ldx rows)
  int ldx, ldw;
  int has_gate;
  int mode;              // requested: 0 full, 1 left factor isometric, 2 right
  int prec;              // contraction: 0 FP64, 1 FP32, 2 BF16 split (FP32 accumulation)
  int tkind;             // 0 None, 1 Const, 2 Cutoff (BondDim, mps.rs:157-166)
  int tchi;
  int cap_rank;          // allocation bound on the new bond dimension
  double teps;
  double2 gate[16];      // 4x4 row-major, qubit k the most significant bit
  const double2* gsrc;   // the gate in device memory instead (generated), or null
  int* hflags;           // the handle's sticky failure flags (kFlag* bits)
  int* eflags;           // the engine's: any handle's flags since the host last looked
};

// The Jacobi step kernels of one launch (all tasks of an update batch).  Task
// t's workspace sits at fixed offsets from these bases, so a step kernel
// computes every address from its arguments and issues its first loads at
// entry.
struct JacobiArgs {
  double2* X;            // X of task t at X + t * xslot (ldx rows)
  double2* Xb;           // or null: a second X (same slots) for the one-launch row split
  double2* W;            // W of task t at W + t * wslot (ldw rows)
  double2* gpart;        // row-split partial Gram matrices, gslot per task:
                         // [pair][slice][PW * PW]
  double* offs;          // [sweep][task] max off-diagonal ratio
  const int* info;       // kInfoLen ints per task
  const int* act;        // or null: the sweep's unconverged tasks (act[0] of them,
                         // indices from act + 1), k_jacobi_step's grid rows
  size_t xslot, wslot, gslot;
  int ldx, ldw;
  int nt;                // tasks in the launch
  int nbe;               // round-robin blocks (even) of the widest task
  int S;                 // workgroups per block pair, a power of two (row split; 1 otherwise)
  int Rh, Ch;            // bounds on every task's R and C
  int inner_every;       // 16 x 16 panel solves rotate the pairs inside each 8-column
                         // block on steps that are multiples of this, capped at a
                         // quarter of the sweep's steps (the pairs across the two
                         // blocks on every step)
  double tol_rot;        // rotate a pair when |g_ij| > tol_rot sqrt(g_ii g_jj)
  double conv_tol;       // converged after a sweep whose max ratio <= conv_tol
  double tiny2;          // columns with squared norm below take no part
};

// The QR preconditioner of one launch batch (tn_qr.hip): the X and W of the
// task in workspace slot t as in JacobiArgs, its factored copy E (same shape
// as X), panel T factors and sort order.
struct QrArgs {
  double2* X;            // X + t * xslot (ldx rows): theta; Y = R^H; finally U Sigma
  double2* W;            // W + t * wslot (ldw rows)
  double2* E;            // E + t * xslot: X with sorted columns, factored in place
  double2* T;            // T + t * tslot: panel jp's 16 x 16 T (row-major) at 256 jp
  double* cn;            // squared column norms, ldw per task
  int* perm0;            // columns by decreasing norm, ldw per task
  int* info;             // kInfoLen ints per task
  const int* map;        // or null: launch task i works in workspace slot map[i] (else slot i)
  size_t xslot, wslot, tslot;
  int ldx, ldw, nt;      // nt: tasks in the launch
};

// One step of the state-vector factorization (mps.rs:448-508 factorize).
// The current right factor q (m x ncols, row-major) is held conjugated in
// its own memory order, which is q^H stored column-major: orientation A
// decomposes X = q^H (ncols x m) in place; orientation B (m > ncols)
// decomposes X = q (m x ncols), a transposed copy.
struct FvArgs {
  double2* X;            // the SVD input, R x C column-major (ld R)
  double2* W;            // C x C accumulated rotation (ld C)
  const double2* src;    // orientation B: conj(q) to transpose into X
  double2* X2;           // the next factor conj(q_new), row-major r x ncols
  double2* gk;           // Gamma[k] (m/2, 2, r)
  const double* lamPrev; // lambda[k-1] or nullptr (k == 0)
  const int* perm;       // sorted column order (k_finish)
  const double* sraw;    // column norms of X W
  const double* fac;     // row scales of the right factor (process_svd)
  double* offs;          // per-sweep maxima, zeroed by k_fv_init
  int* info;             // kInfoLen ints, written by k_fv_init
  int m, ncols, R, C, orient, rank;
};

// One entropy read-out: lambda of a bond, where its length is, Renyi order a
// (1: von Neumann), result slot.
struct EntDesc {
  const double* lam;
  const int* dim;
  double a;
  int slot;
};

// One site measurement or probability query (mps.rs:1027-1134, 1427-1465).
struct MeasDesc {
  double2* g;            // Gamma[k]
  const double* lamL;    // lambda[k-1] or nullptr
  const double* lamR;    // lambda[k] or nullptr
  const int* dims;       // &dims[k]: chiL = dims[0], chiR = dims[1]
  int forced;            // 0/1 forced outcome, -1 sample with r, -2 query only
  int slot;              // index into the result arrays
  double r;              // uniform draw for sampling
  int* out_outcome;
  double* out_prob;      // outcome probability
  double* out_probs;     // both probabilities, 2 per slot
  int* out_flag;         // 1 zero probability, 2 draw out of range
  int* out_dims;         // or null: chiL, chiR of descriptor i at out_dims[2 i]
};

}  // namespace tnmps
