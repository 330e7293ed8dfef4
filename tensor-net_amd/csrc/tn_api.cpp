// Host driver and C ABI of the MI355X Schmidt-form MPS update path.
//
// The driver restates the control flow of tensor-net's MPS methods
// (tensor-net/lib/mps.rs) -- which bonds are refactored in which order, when
// sites are measured, what is a no-op -- and enqueues the device work for
// each step on the engine's HIP stream.  All numerical work happens in the
// kernels of tn_kernels.hip; Gamma/lambda never leave the device except when
// the caller asks for them.
//
// An engine owns a stream, the update workspace and the descriptor staging.
// Handles bound to one engine can be advanced in lock step: the updates and
// measurements that different handles need at the same point of their
// schedules run in shared launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "tn_device.h"
#include "tn_rng.h"
#include "tn_kernels.h"
#include "tnmps.h"

using namespace tnmps;

namespace {

thread_local std::string g_last_error;

#define TN_HIP(expr)                                                              \
  do {                                                                            \
    hipError_t e__ = (expr);                                                      \
    if (e__ != hipSuccess) {                                                      \
      g_last_error = std::string(#expr) + ": " + hipGetErrorString(e__);          \
      return TN_DEVICE;                                                           \
    }                                                                             \
  } while (0)

#define TN_TRY(expr)                \
  do {                              \
    tn_status s__ = (expr);         \
    if (s__ != TN_OK) return s__;   \
  } while (0)

double env_double(const char* name, double dflt) {
  const char* v = std::getenv(name);
  if (!v || !*v) return dflt;
  return std::atof(v);
}

// Jacobi tolerances.  A pair of columns is rotated when |g_ij| exceeds
// tol_rot * sqrt(g_ii g_jj); the SVD has converged after the first sweep whose
// largest such ratio is at most conv_tol.  That sweep still rotates every pair
// above tol_rot, and Jacobi converges quadratically, so the residual after it
// is of order conv_tol^2 = 1e-14 (the sweeps measured here go 2e-3 -> 2e-8
// -> 1e-14 and 1e-4 -> 3e-9: a sweep that starts at 1e-7 ends near 1e-14,
// so a confirming sweep would only measure that).
// Both are read from the environment when an engine is made.
// A column whose norm is below tiny_norm (relative to the unit-norm SVD
// input) is numerically null: it takes part in no rotation and in no
// convergence measure.  Below f64::EPSILON, such a column's value is dropped
// by the rank rule (mps.rs:255-262) whatever further rotations would make of
// it, and a rotation only shrinks the smaller column of its pair.
//
// A theta contracted in FP32 (or BF16 split) carries relative round-off near
// 6e-8 (1.5e-5) in every entry; orthogonalising its columns to 1e-16 buys
// nothing.  Such a task's SVD has converged after a sweep whose largest ratio
// is at most conv_tol_f32 (conv_tol_bf16), leaving a residual near its square
// (1e-8, 1e-6), below theta's own round-off.
// The residual only tilts the new left factor's columns from orthogonal by
// that much: X W = U~ exactly with W unitary, so the split still writes theta
// exactly.  A launch batch holds tasks of one precision (run_tasks).
struct Tolerances {
  double tol_rot = 1e-14;
  double conv_tol = 1e-7;
  double conv_tol_f32 = 1e-4;
  double conv_tol_bf16 = 1e-3;
  double tiny_norm = 1.5e-16;
  void read_env() {
    tol_rot = env_double("TNMPS_TOL_ROT", tol_rot);
    conv_tol = env_double("TNMPS_CONV_TOL", conv_tol);
    conv_tol_f32 = env_double("TNMPS_CONV_TOL_F32", conv_tol_f32);
    conv_tol_bf16 = env_double("TNMPS_CONV_TOL_BF16", conv_tol_bf16);
    tiny_norm = env_double("TNMPS_TINY_NORM", tiny_norm);
  }
  double for_prec(int prec) const {
    return prec == TN_PREC_F32 ? conv_tol_f32 : prec == TN_PREC_BF16X3 ? conv_tol_bf16 : conv_tol;
  }
};

void hfree(void* p) {
  if (p) (void)hipFree(p);
}
void hhfree(void* p) {
  if (p) (void)hipHostFree(p);
}

constexpr int kRing = 4;  // staging buffers in flight per descriptor kind
// tn_matrix_part: at most 12 sites (a 4096 x 4096 matrix) and at most 2^28
// complex entries (4 GiB) in the contracted block of the range
constexpr size_t kMatrixPartMaxSites = 12;
constexpr size_t kMatrixPartMaxElems = (size_t)1 << 28;

size_t exact_bond_max(size_t n, size_t b) {
  // min(2^{b+1}, 2^{n-b-1}) without overflow
  const size_t l = b + 1, r = n - b - 1;
  const size_t e = std::min(l, r);
  if (e >= 40) return (size_t)1 << 40;
  return (size_t)1 << e;
}

// Pinned host staging + device copy of a descriptor array, in a ring so the
// host can prepare the next launch while earlier ones are still queued.
template <typename T>
struct DescRing {
  int cap = 0;
  int pos = 0;
  T* h[kRing] = {};
  T* d[kRing] = {};
  hipEvent_t ev[kRing] = {};

  void release() {
    for (int i = 0; i < kRing; ++i) {
      hhfree(h[i]);
      hfree(d[i]);
      if (ev[i]) (void)hipEventDestroy(ev[i]);
      h[i] = nullptr;
      d[i] = nullptr;
      ev[i] = nullptr;
    }
    cap = 0;
  }
  tn_status ensure(int need, hipStream_t s) {
    if (need <= cap) return TN_OK;
    TN_HIP(hipStreamSynchronize(s));
    release();
    const int c = std::max(need, 64);
    for (int i = 0; i < kRing; ++i) {
      TN_HIP(hipHostMalloc(&h[i], sizeof(T) * c, hipHostMallocDefault));
      TN_HIP(hipMalloc(&d[i], sizeof(T) * c));
      TN_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
    }
    cap = c;
    return TN_OK;
  }
  // Host buffer for the next launch (waits until its previous copy is done).
  tn_status next(T** host, int* slot) {
    const int i = pos;
    pos = (pos + 1) % kRing;
    TN_HIP(hipEventSynchronize(ev[i]));
    *host = h[i];
    *slot = i;
    return TN_OK;
  }
  tn_status upload(int slot, int count, hipStream_t s) {
    TN_HIP(hipMemcpyAsync(d[slot], h[slot], sizeof(T) * count, hipMemcpyHostToDevice, s));
    TN_HIP(hipEventRecord(ev[slot], s));
    return TN_OK;
  }
};

using EntReq = EntDesc;

// ---- static scratch of tn_apply_u2_noalloc -------------------------------
// One two-site update's workspace with theta up to kScratchCols x
// kScratchCols (bonds next to the updated one up to kScratchCols / 2), laid
// out as the engine lays out one slot: X and W (kScratchCols^2 complex
// each), the per-column arrays, partial norms, per-sweep maxima, info, the
// active-task list, the row-split partial Gram matrices (both block widths),
// the QR preconditioner's factored copy of X, panel T factors and sort
// order, and the task's descriptor.  Each device holds one copy for the life of the
// process; its address on a device comes from hipGetSymbolAddress.
constexpr int kScratchCols = 512;
constexpr size_t scratch_align(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kScrX = 0;
constexpr size_t kScrW = kScrX + scratch_align(sizeof(double2) * kScratchCols * kScratchCols);
constexpr size_t kScrSraw = kScrW + scratch_align(sizeof(double2) * kScratchCols * kScratchCols);
constexpr size_t kScrSval = kScrSraw + scratch_align(sizeof(double) * kScratchCols);
constexpr size_t kScrFac = kScrSval + scratch_align(sizeof(double) * kScratchCols);
constexpr size_t kScrPerm = kScrFac + scratch_align(sizeof(double) * kScratchCols);
constexpr size_t kScrPart = kScrPerm + scratch_align(sizeof(int) * kScratchCols);
constexpr size_t kScrOffs = kScrPart + scratch_align(sizeof(double) * 2 * kNPart);
constexpr size_t kScrInfo = kScrOffs + scratch_align(sizeof(double) * kMaxSweeps);
constexpr size_t kScrAct = kScrInfo + scratch_align(sizeof(int) * kInfoLen);
constexpr size_t kScrGpart = kScrAct + scratch_align(sizeof(int) * 2);
// partial Gram matrices of one task: (pairs + 1) x kMaxSplit slices of
// (2 bw)^2 values, for the wider of the two block widths
constexpr size_t gpart_elems(int cols, int bw) {
  return ((size_t)(cols + bw - 1) / bw / 2 + 1) * kMaxSplit * (size_t)(4 * bw * bw);
}
constexpr size_t kScrGpartElems = gpart_elems(kScratchCols, kBW) > gpart_elems(kScratchCols, kBWWide)
                                      ? gpart_elems(kScratchCols, kBW)
                                      : gpart_elems(kScratchCols, kBWWide);
constexpr size_t kScrQrE = kScrGpart + scratch_align(sizeof(double2) * kScrGpartElems);
constexpr size_t kScrQrTElems = (size_t)(kScratchCols / kQrNb) * kQrNb * kQrNb;
constexpr size_t kScrQrT = kScrQrE + scratch_align(sizeof(double2) * kScratchCols * kScratchCols);
constexpr size_t kScrQrPerm = kScrQrT + scratch_align(sizeof(double2) * kScrQrTElems);
constexpr size_t kScrXb = kScrQrPerm + scratch_align(sizeof(int) * kScratchCols);
constexpr size_t kScrDesc = kScrXb + scratch_align(sizeof(double2) * kScratchCols * kScratchCols);
constexpr size_t kScratchBytes = kScrDesc + scratch_align(sizeof(UpdDesc));

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// This device's address of `scratch`, resolved once per device with that
// device current.  (The array's name in host code is a host-side stand-in:
// only the address returned here is a device address.)
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

tn_status scratch_address(int dev, void** out) {
  *out = nullptr;
  if (dev < 0 || dev >= kMaxDevices) return TN_DEVICE;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    TN_HIP(hipGetDevice(&cur));
    if (cur != dev) TN_HIP(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t er = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (er != hipSuccess) {
      g_last_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(er);
      return TN_DEVICE;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TN_OK;
}

}  // namespace

// Device-time accounting of one update batch or measurement launch.  Its
// events and a pinned copy of the batch's info are read back lazily (when
// they have completed, or when the counters are asked for), so timing adds
// no host synchronisation to the stream.
struct PerfRec {
  // theta begin/end, sweeps begin/end, post begin/end, split begin, renorm
  // begin, contraction end (QR begin), QR end (winit begin); ev[3]..ev[4]
  // spans the QR's last kernel
  hipEvent_t ev[10] = {};
  bool qr = false;        // the batch ran the QR preconditioner
  int* h_info = nullptr;  // pinned copy of the batch's info (update batch) or the
                          // measured sites' chiL, chiR (measurement launch)
  int cap = 0;            // tasks (or sites) h_info holds
  std::vector<char> proj; // measurement launch: the site is projected
  std::vector<char> gate; // has_gate per task
  std::vector<int> slots; // pool records: task t's info sits at slot slots[t] of h_info
  int kind = 0;           // 0 update batch, 1 measurement launch (ev[0], ev[1];
                          // ev[2] once the sites' dimensions are on the host),
                          // 2 one pool sweep (ev[0], ev[1])
  int nt = 0, n_updates = 0, sweeps = 0;
  size_t launches = 0;
  bool pending = false;
};
constexpr int kPerfRing = 16;

// --------------------------------------------------------------------------
// Engine: stream, update workspace, descriptor staging, result buffers.
// --------------------------------------------------------------------------
struct tn_engine {
  int dev = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  size_t chi_max = 0;   // largest bond dimension of any handle served
  int ldx = 0, ldw = 0; // row capacities of X (R) and W (C)
  int slots = 0;        // update tasks the workspace holds
  int max_slots = 0;    // tasks per launch batch (larger batches are chunked)
  double2* X = nullptr;
  double2* W = nullptr;
  double* sraw = nullptr;
  double* sval = nullptr;
  double* fac = nullptr;
  int* perm = nullptr;
  double* part = nullptr;
  double2* gpart = nullptr;
  size_t gpart_per = 0;  // elements of gpart per slot
  int gpart_slots = 0;   // slots gpart covers: the row split (S > 1) only runs
                         // for batches of at most `cus` tasks
  double* offs = nullptr;
  int* info = nullptr;
  int* act = nullptr;    // active-task list of the current sweep: count, then tasks
  // QR preconditioner (tn_qr.hip): factored copy of X, panel T factors and
  // sort order per slot
  double2* qrE = nullptr;
  double2* qrT = nullptr;
  int* qrPerm = nullptr;
  size_t qr_tslot = 0;   // elements of qrT per slot
  // second X of the first xb_slots slots: the one-launch row-split step
  // ping-pongs X between the two (a lone SVD or a few)
  double2* Xb = nullptr;
  int xb_slots = 0;
  bool split_fused = true;  // one-launch row-split steps (TNMPS_SPLIT_FUSED)
  // least rows per slice of the one-launch step, whose split doubles past
  // split_rows while its slices keep at least this many (TNMPS_FUSED_ROWS)
  int fused_rows = 32;
  // the 16 x 16 panel solves of FP64 one-launch row-split steps rotate the
  // pairs inside each 8-column block on every inner_every-th step only,
  // and on at least four steps of every sweep (TNMPS_INNER_EVERY; 1: every step)
  int inner_every = 8;
  int qr_min = 32;       // least C for which a batch takes the QR path (TNMPS_QR_MIN)
  int cus = 0;           // compute units of the device
  int list_min = 64;     // tasks from which the step kernel runs over act (TNMPS_LIST_MIN)
  // SVD task pool (run_pool): handles from which a batch's measurement
  // schedules run as a pool (TNMPS_POOL_MIN; 0, the default, keeps the
  // lock-step path: on the measured workloads the two ran at the same rate),
  // its pinned active list and slot maps and their device copies
  int pool_min = 0;
  int pool_cap = 0;
  int* h_pact = nullptr;
  int* h_pmap[2] = {};
  int* d_pmap[2] = {};
  int split_max = kMaxSplit;  // cap on workgroups per block pair (TNMPS_SPLIT_MAX)
  int split_rows = 128;       // the split doubles while rows >= split_rows x S (TNMPS_SPLIT_ROWS)
  int bw_env = 0;             // forced Jacobi block width (TNMPS_BW), 0 = automatic
  std::vector<tn_tune> tune;  // launch shapes per C bucket (tn_engine_set_tuning)
  bool factored = true;       // factored updates where a factor is isometric (TNMPS_FACTORED)
  Tolerances tol;             // Jacobi tolerances and inner-solve options
  bool trace_sweeps = false;  // print each sweep's largest off-diagonal (TNMPS_TRACE_SWEEPS)
  DescRing<UpdDesc> upd;
  DescRing<MeasDesc> meas;
  DescRing<EntDesc> ents;
  int mres = 0;  // measurement result slots
  double* d_mpart = nullptr;   // per-site partial probabilities (2 kMeasParts each)
  int* d_out = nullptr;
  double* d_prob = nullptr;
  double* d_probs = nullptr;
  int* d_flag = nullptr;
  int* d_mdims = nullptr;      // measured sites' chiL, chiR (timing only)
  int n_mdims = 0;
  double* d_scalar = nullptr;  // scalar results (entropies)
  int nscalar = 0;
  double2* gen_u = nullptr;    // generated gates, 16 entries each
  uint64_t* gen_ids = nullptr; // their counter words, 3 each
  size_t gen_cap = 0;
  bool timing = false;
  int max_sweeps = kMaxSweeps;  // Jacobi sweep limit (TNMPS_MAX_SWEEPS, 1..kMaxSweeps)
  // Sticky failure flags (kFlag* bits) of the updates enqueued since the host
  // last looked: the device word (rank clamps, set by k_finish), its pinned
  // copy, and the host's own (sweep limit reached).
  int* d_eflags = nullptr;
  int* h_eflags = nullptr;
  int host_flags = 0;
  unsigned char* scratch_ptr = nullptr;  // this device's `scratch` (tn_apply_u2_noalloc)
  double perf_host_launch_ms = 0.0;  // host time spent enqueueing Jacobi steps
  hipEvent_t sev[kMaxSweeps] = {};  // per-sweep completion (convergence readback)
  double* h_offs = nullptr;         // pinned copy of offs
  PerfRec prec[kPerfRing];          // timing records, oldest at prec_next
  int prec_next = 0;
  tn_perf perf{};
  unsigned long long updates = 0, sweeps = 0;
  int users = 0;               // handles bound to this engine
  bool auto_free = false;      // freed with its last handle

  ~tn_engine() { release(); }

  // the first row whose bucket holds C, else the last row
  const tn_tune& tune_for(size_t C) const {
    for (const tn_tune& r : tune)
      if ((size_t)r.c_max >= C) return r;
    return tune.back();
  }

  void release_ws() {
    hfree(X); hfree(W); hfree(sraw); hfree(sval); hfree(fac); hfree(perm);
    hfree(part); hfree(offs); hfree(info); hfree(gpart); hfree(act);
    hfree(qrE); hfree(qrT); hfree(qrPerm); hfree(Xb);
    qrE = qrT = Xb = nullptr;
    qrPerm = nullptr;
    xb_slots = 0;
    hhfree(h_offs);
    h_offs = nullptr;
    gpart = nullptr;
    gpart_slots = 0;
    X = W = nullptr;
    sraw = sval = fac = part = offs = nullptr;
    perm = info = act = nullptr;
    slots = 0;
  }
  void release_res() {
    hfree(d_out); hfree(d_prob); hfree(d_probs); hfree(d_flag); hfree(d_mdims); hfree(d_mpart);
    d_mpart = nullptr;
    d_out = nullptr; d_prob = d_probs = nullptr; d_flag = nullptr; d_mdims = nullptr;
    n_mdims = 0;
    mres = 0;
  }
  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    release_ws();
    release_res();
    upd.release();
    meas.release();
    ents.release();
    hfree(d_scalar); d_scalar = nullptr; nscalar = 0;
    hhfree(h_pact); h_pact = nullptr;
    for (int i = 0; i < 2; ++i) {
      hhfree(h_pmap[i]); hfree(d_pmap[i]);
      h_pmap[i] = nullptr; d_pmap[i] = nullptr;
    }
    pool_cap = 0;
    hfree(d_eflags); d_eflags = nullptr;
    hhfree(h_eflags); h_eflags = nullptr;
    hfree(gen_u); hfree(gen_ids); gen_u = nullptr; gen_ids = nullptr; gen_cap = 0;
    for (PerfRec& r : prec) {
      for (hipEvent_t& ev : r.ev)
        if (ev) (void)hipEventDestroy(ev), ev = nullptr;
      hhfree(r.h_info);
      r.h_info = nullptr;
      r.cap = 0;
      r.pending = false;
    }
    for (hipEvent_t& ev : sev)
      if (ev) (void)hipEventDestroy(ev), ev = nullptr;
    if (own_stream && stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }

  tn_status init(void* s, size_t chi) {
    TN_HIP(hipGetDevice(&dev));
    if (s) {
      stream = (hipStream_t)s;
      own_stream = false;
    } else {
      TN_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
      own_stream = true;
    }
    chi_max = chi;
    ldx = ldw = (int)(2 * chi);
    for (PerfRec& r : prec)
      for (hipEvent_t& ev : r.ev) TN_HIP(hipEventCreate(&ev));
    for (hipEvent_t& ev : sev) TN_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    hipDeviceProp_t prop;
    TN_HIP(hipGetDeviceProperties(&prop, dev));
    cus = prop.multiProcessorCount;
    split_max = std::max(1, std::min(kMaxSplit, (int)env_double("TNMPS_SPLIT_MAX", kMaxSplit)));
    split_rows = std::max(64, (int)env_double("TNMPS_SPLIT_ROWS", 128));
    split_fused = env_double("TNMPS_SPLIT_FUSED", 1) != 0;
    fused_rows = std::max(16, (int)env_double("TNMPS_FUSED_ROWS", 32));
    inner_every = std::max(1, (int)env_double("TNMPS_INNER_EVERY", 8));
    list_min = std::max(1, (int)env_double("TNMPS_LIST_MIN", 64));
    pool_min = std::max(0, (int)env_double("TNMPS_POOL_MIN", 0));
    size_t gp = 0;
    for (int bw : {kBW, kBWWide}) {
      const size_t nbmax = (size_t)(ldw + bw - 1) / bw;
      gp = std::max(gp, (nbmax / 2 + 1) * kMaxSplit * (size_t)(4 * bw * bw));
    }
    gpart_per = gp;
    bw_env = (int)env_double("TNMPS_BW", 0);
    // default table: one row per bucket; tiles of 32 keep small thetas on
    // more workgroups, 64 halves the LDS traffic per flop of large ones
    tune.clear();
    for (int cm : {16, 32, 64, 128, 256, 512, 1024}) {
      tn_tune r{};
      r.c_max = cm;
      r.theta_tile = (int)env_double("TNMPS_THETA_TILE", cm <= 128 ? 32 : 64) == 32 ? 32 : 64;
      r.bw = bw_env == kBWWide ? kBWWide : kBW;
      r.split_max = split_max;
      r.split_rows = split_rows;
      r.ahead = std::max(1, std::min(3, (int)env_double("TNMPS_AHEAD", 2)));
      r.qr = env_double("TNMPS_QR", 1) != 0 ? 1 : 0;
      tune.push_back(r);
    }
    factored = env_double("TNMPS_FACTORED", 1) != 0;
    qr_min = std::max(kQrNb, (int)env_double("TNMPS_QR_MIN", 32));
    qr_tslot = (size_t)((ldw + kQrNb - 1) / kQrNb) * kQrNb * kQrNb;
    max_sweeps = std::max(1, std::min(kMaxSweeps, (int)env_double("TNMPS_MAX_SWEEPS", kMaxSweeps)));
    {
      void* sp = nullptr;
      TN_TRY(scratch_address(dev, &sp));
      scratch_ptr = (unsigned char*)sp;
    }
    TN_HIP(hipMalloc(&d_eflags, sizeof(int)));
    TN_HIP(hipHostMalloc(&h_eflags, sizeof(int), hipHostMallocDefault));
    *h_eflags = 0;
    TN_HIP(hipMemcpyAsync(d_eflags, h_eflags, sizeof(int), hipMemcpyHostToDevice, stream));
    TN_HIP(hipStreamSynchronize(stream));
    tol.read_env();
    trace_sweeps = env_double("TNMPS_TRACE_SWEEPS", 0) != 0;
    // workspace budget: TNMPS_WS_BYTES (default 24 GiB) bounds the tasks per launch
    const double budget = env_double("TNMPS_WS_BYTES", 24.0 * (1ull << 30));
    const double per = 16.0 * (2.0 * ldx * ldw + (double)ldw * ldw + (double)qr_tslot) +
                       68.0 * ldw + 4096.0;
    max_slots = (int)std::max(1.0, std::min(8192.0, budget / per));
    return TN_OK;
  }

  static float span(const PerfRec& r, int a, int b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.ev[a], r.ev[b]) != hipSuccess) return 0.f;
    return ms;
  }

  // The event a record's last span ends with.
  static int last_ev(const PerfRec& r) { return r.kind == 1 ? 2 : r.kind == 2 ? 1 : 5; }

  // Add a completed record to the counters.
  void perf_process(PerfRec& r) {
    if (r.kind == 2) {
      perf.ms_jacobi += span(r, 0, 1);
      perf.jacobi_launches += r.launches;
      perf.sweeps += 1;
    } else if (r.kind == 1) {
      perf.ms_measure += span(r, 0, 1);
      for (int i = 0; i < r.nt; ++i) {
        const double cells = (double)r.h_info[2 * i] * (double)r.h_info[2 * i + 1];
        perf.bytes_measure += 32.0 * cells * (r.proj[i] ? 2.0 : 1.0);
      }
      perf.measure_sites += (size_t)r.nt;
    } else {
      perf.ms_theta += span(r, 0, 8) + span(r, 9, 1);
      perf.ms_contract += span(r, 0, 8);
      perf.ms_qr += span(r, 8, 9) + span(r, 3, 4);
      perf.ms_jacobi += span(r, 2, 3);
      perf.ms_post += span(r, 4, 5);
      perf.ms_split += span(r, 6, 7);
      perf.ms_renorm += span(r, 7, 5);
      perf.jacobi_launches += r.launches;
      perf.updates += r.n_updates;
      perf.sweeps += r.sweeps;
      for (int t = 0; t < r.nt; ++t) {
        // flops of the SVD posed (R x C; LAPACK-style count of a complex SVD
        // with both factors) plus the contraction or the factor product
        const int* inf = r.h_info + (size_t)(r.slots.empty() ? t : r.slots[t]) * kInfoLen;
        const double m = inf[kInfoM], n = inf[kInfoN], K = inf[kInfoChiM];
        const double M = inf[kInfoR], N = inf[kInfoC], rank = inf[kInfoRank];
        const double fsvd = 4.0 * (4.0 * M * M * N + 8.0 * M * N * N + 9.0 * N * N * N);
        perf.svd_flops += fsvd;
        const int mode = inf[kInfoMode];
        if (mode != 0) perf.factored++;
        const double fpre = mode == 0 ? 8.0 * m * n * K + (r.gate[t] ? 32.0 * m * n : 0.0)
                                      : 8.0 * (mode == 1 ? m : n) * K * rank;
        perf.alg_flops += fpre + fsvd;
        if (mode == 0) perf.contract_flops += fpre;
        perf.bytes_split += 32.0 * rank * (m + n);
        perf.bytes_renorm += 32.0 * rank * (m + n);
      }
    }
    r.pending = false;
  }

  // Process pending records oldest first; without `block`, stop at the first
  // one whose events have not completed.
  tn_status perf_drain(bool block) {
    for (int i = 0; i < kPerfRing; ++i) {
      PerfRec& r = prec[(prec_next + i) % kPerfRing];
      if (!r.pending) continue;
      const hipEvent_t last = r.ev[last_ev(r)];
      if (block) {
        TN_HIP(hipEventSynchronize(last));
      } else if (hipEventQuery(last) != hipSuccess) {
        break;
      }
      perf_process(r);
    }
    return TN_OK;
  }

  // The next record, processed first if still pending, with room for nt tasks.
  tn_status perf_begin(int kind, int nt, PerfRec** out) {
    PerfRec& r = prec[prec_next];
    if (r.pending) {
      TN_HIP(hipEventSynchronize(r.ev[last_ev(r)]));
      perf_process(r);
    }
    if (nt > r.cap) {
      hhfree(r.h_info);
      r.h_info = nullptr;
      r.cap = 0;
      TN_HIP(hipHostMalloc(&r.h_info, sizeof(int) * kInfoLen * nt, hipHostMallocDefault));
      r.cap = nt;
    }
    r.kind = kind;
    r.nt = nt;
    r.slots.clear();
    r.launches = 0;
    r.n_updates = r.sweeps = 0;
    prec_next = (prec_next + 1) % kPerfRing;
    *out = &r;
    return TN_OK;
  }

  tn_status ensure_slots(int need) {
    if (need <= slots) return TN_OK;
    TN_HIP(hipStreamSynchronize(stream));
    release_ws();
    const int ns = need;
    const size_t xe = (size_t)ldx * ldw, we = (size_t)ldw * ldw;
    TN_HIP(hipMalloc(&X, sizeof(double2) * xe * ns));
    TN_HIP(hipMalloc(&W, sizeof(double2) * we * ns));
    TN_HIP(hipMalloc(&sraw, sizeof(double) * ldw * ns));
    TN_HIP(hipMalloc(&sval, sizeof(double) * ldw * ns));
    TN_HIP(hipMalloc(&fac, sizeof(double) * ldw * ns));
    TN_HIP(hipMalloc(&perm, sizeof(int) * ldw * ns));
    TN_HIP(hipMalloc(&part, sizeof(double) * 2 * kNPart * ns));
    TN_HIP(hipMalloc(&offs, sizeof(double) * kMaxSweeps * ns));
    TN_HIP(hipHostMalloc(&h_offs, sizeof(double) * kMaxSweeps * ns, hipHostMallocDefault));
    TN_HIP(hipMalloc(&info, sizeof(int) * kInfoLen * ns));
    gpart_slots = std::min(ns, std::max(cus, 1));
    TN_HIP(hipMalloc(&gpart, sizeof(double2) * gpart_per * gpart_slots));
    TN_HIP(hipMalloc(&act, sizeof(int) * (ns + 1)));
    TN_HIP(hipMalloc(&qrE, sizeof(double2) * xe * ns));
    TN_HIP(hipMalloc(&qrT, sizeof(double2) * qr_tslot * ns));
    TN_HIP(hipMalloc(&qrPerm, sizeof(int) * ldw * ns));
    xb_slots = std::min(ns, kFusedSlots);
    TN_HIP(hipMalloc(&Xb, sizeof(double2) * xe * xb_slots));
    slots = ns;
    return upd.ensure(ns, stream);
  }

  tn_status ensure_pool(int need) {
    if (need <= pool_cap) return TN_OK;
    TN_HIP(hipStreamSynchronize(stream));
    hhfree(h_pact);
    h_pact = nullptr;
    for (int i = 0; i < 2; ++i) {
      hhfree(h_pmap[i]); hfree(d_pmap[i]);
      h_pmap[i] = nullptr; d_pmap[i] = nullptr;
    }
    pool_cap = 0;
    TN_HIP(hipHostMalloc(&h_pact, sizeof(int) * (need + 1), hipHostMallocDefault));
    for (int i = 0; i < 2; ++i) {
      TN_HIP(hipHostMalloc(&h_pmap[i], sizeof(int) * need, hipHostMallocDefault));
      TN_HIP(hipMalloc(&d_pmap[i], sizeof(int) * need));
    }
    pool_cap = need;
    return TN_OK;
  }

  tn_status ensure_mdims(int need) {
    if (need <= n_mdims) return TN_OK;
    TN_HIP(hipStreamSynchronize(stream));
    hfree(d_mdims);
    d_mdims = nullptr;
    n_mdims = 0;
    const int ns = std::max(need, 256);
    TN_HIP(hipMalloc(&d_mdims, sizeof(int) * 2 * ns));
    n_mdims = ns;
    return TN_OK;
  }

  tn_status ensure_results(int need) {
    if (need <= mres) return TN_OK;
    TN_HIP(hipStreamSynchronize(stream));
    release_res();
    const int ns = std::max(need, 256);
    TN_HIP(hipMalloc(&d_out, sizeof(int) * ns));
    TN_HIP(hipMalloc(&d_prob, sizeof(double) * ns));
    TN_HIP(hipMalloc(&d_probs, sizeof(double) * 2 * ns));
    TN_HIP(hipMalloc(&d_flag, sizeof(int) * ns));
    TN_HIP(hipMalloc(&d_mpart, sizeof(double) * 2 * kMeasParts * ns));
    mres = ns;
    return TN_OK;
  }

  // Room for `need` generated gates (the stream is drained before a resize,
  // so no pending launch reads a freed buffer).
  tn_status ensure_gen(size_t need) {
    if (need <= gen_cap) return TN_OK;
    TN_HIP(hipStreamSynchronize(stream));
    hfree(gen_u);
    hfree(gen_ids);
    gen_u = nullptr;
    gen_ids = nullptr;
    gen_cap = 0;
    const size_t cap = std::max<size_t>(need, 1024);
    TN_HIP(hipMalloc(&gen_u, sizeof(double2) * 16 * cap));
    TN_HIP(hipMalloc(&gen_ids, sizeof(uint64_t) * 3 * cap));
    gen_cap = cap;
    return TN_OK;
  }

  tn_status ensure_scalars(int need) {
    if (need <= nscalar) return TN_OK;
    TN_HIP(hipStreamSynchronize(stream));
    hfree(d_scalar);
    d_scalar = nullptr;
    const int ns = std::max(need, 256);
    TN_HIP(hipMalloc(&d_scalar, sizeof(double) * ns));
    nscalar = ns;
    return TN_OK;
  }
};

// --------------------------------------------------------------------------
// MPS handle
// --------------------------------------------------------------------------
struct tn_mps {
  tn_engine* eng = nullptr;
  size_t n = 0;
  tn_trunc trunc{};
  size_t cap = 0;
  int prec = TN_PREC_F64;
  std::vector<size_t> bmax;     // allocation bound per bond
  std::vector<size_t> goff;     // element offset of each site slot
  std::vector<size_t> loff;     // element offset of each bond slot
  double2* gam = nullptr;
  double* lam = nullptr;
  int* dims = nullptr;
  std::vector<size_t> ub;       // host-side upper bound of each bond dimension
  // Per site: kFreshLeft when lambda[k-1] Gamma[k] has orthonormal columns
  // (the site was last written as the left factor of an update at bond k),
  // kFreshRight when Gamma[k] lambda[k] has orthonormal rows (last written as
  // the right factor of an update at bond k-1).  A projection or a one-site
  // operator clears both; the product state has both.
  std::vector<unsigned char> fresh;
  size_t gam_elems = 0, lam_elems = 0;
  // Sticky failure flags (kFlag* bits) of this handle's updates: the device
  // word sits after the bond dimensions (dims[n + 1]); collected ones are
  // kept here until tn_take_flags.
  int* dflags() const { return dims + n + 1; }
  int flags = 0;

  ~tn_mps() {
    if (eng) (void)hipStreamSynchronize(eng->stream);
    hfree(gam);
    hfree(lam);
    hfree(dims);
    if (eng) {
      eng->users--;
      if (eng->auto_free && eng->users <= 0) delete eng;
    }
  }
  size_t bL(size_t k) const { return k == 0 ? 1 : bmax[k - 1]; }
  size_t bR(size_t k) const { return k + 1 >= n ? 1 : bmax[k]; }
  size_t uL(size_t k) const { return k == 0 ? 1 : ub[k - 1]; }
  size_t uR(size_t k) const { return k + 1 >= n ? 1 : ub[k]; }
};

namespace {

size_t rank_bound(const tn_mps* h) {
  if (h->trunc.kind == TN_TRUNC_CONST) return std::max<size_t>(h->trunc.chi, 1);
  return (size_t)-1;
}

constexpr unsigned char kFreshLeft = 1, kFreshRight = 2;

// Factored mode of a refactor of bond k (no gate): 1 when site k is
// left-fresh, 2 when site k+1 is right-fresh, the one with fewer SVD rows
// when both are; 0 for the full theta.
int factor_mode(const tn_engine* e, const tn_mps* h, size_t k) {
  if (!e->factored) return 0;
  const bool fl = (h->fresh[k] & kFreshLeft) != 0;
  const bool fr = (h->fresh[k + 1] & kFreshRight) != 0;
  if (fl && fr) return h->uR(k + 1) <= h->uL(k) ? 1 : 2;
  return fl ? 1 : (fr ? 2 : 0);
}

// Gauge prepass of a refactor of bond k with neither site fresh but
// K < min(m, n): theta = A M with A = lambda Gamma[k] lambda[k] (m x K) and
// M = Gamma[k+1] lambda[k+1] (K x n).  Mode 2 on A (or mode 1 on M^H, the
// one with fewer rows) is a valid gauge change even when the other factor is
// not an isometry: it writes the same theta up to a positive scale, with
// site k left-fresh (site k+1 right-fresh).  The refactor that follows is
// then factored too: two SVDs of K columns instead of one of a
// rank-deficient min(m, n)-column theta.  0: no prepass.
int prepass_mode(const tn_engine* e, const tn_mps* h, size_t k) {
  if (!e->factored) return 0;
  if ((h->fresh[k] & kFreshLeft) || (h->fresh[k + 1] & kFreshRight)) return 0;
  const size_t m = 2 * h->uL(k), n = 2 * h->uR(k + 1);
  if (!(h->ub[k] < std::min(m, n))) return 0;
  return m <= n ? 2 : 1;
}

// One two-site update of one handle.
struct Task {
  tn_mps* h;
  size_t k;
  bool has_gate;
  int pre = 0;  // gauge prepass in this mode (not an update of the reference)
  double gate[32];
  const double2* gdev = nullptr;  // the gate in device memory (generated), else gate[]
};

// One site measurement (or probability query with forced == -2).
struct MTask {
  tn_mps* h;
  size_t k;
  int forced;
  double r;
  int slot;
};

// Device workspace of one launch batch: task t's X (ldx x ldw), W (ldw x
// ldw), column norms, values, scales, order, partial norms, per-sweep maxima
// (stride nt) and info, at fixed offsets from these bases; the row-split
// partial Gram matrices of the first gpart_slots tasks; the active-task
// list; the host copy of the per-sweep maxima; and, when not null, the
// device copy of the batch's descriptors (else the engine's ring holds it).
struct Workspace {
  double2* X = nullptr;
  double2* W = nullptr;
  double* sraw = nullptr;
  double* sval = nullptr;
  double* fac = nullptr;
  int* perm = nullptr;
  double* part = nullptr;
  double2* gpart = nullptr;
  size_t gpart_per = 0;
  int gpart_slots = 0;
  double* offs = nullptr;
  int* info = nullptr;
  int* act = nullptr;
  double* h_offs = nullptr;
  double2* qrE = nullptr;   // or null: no QR preconditioner in this workspace
  double2* qrT = nullptr;
  int* qrPerm = nullptr;
  size_t qr_tslot = 0;
  double2* Xb = nullptr;    // or null: row-split steps take two launches
  int xb_slots = 0;
  int ldx = 0, ldw = 0;
  UpdDesc* desc = nullptr;
};

Workspace engine_workspace(const tn_engine* e) {
  Workspace w;
  w.X = e->X; w.W = e->W; w.sraw = e->sraw; w.sval = e->sval; w.fac = e->fac;
  w.perm = e->perm; w.part = e->part; w.gpart = e->gpart; w.gpart_per = e->gpart_per;
  w.gpart_slots = e->gpart_slots; w.offs = e->offs; w.info = e->info; w.act = e->act;
  w.h_offs = e->h_offs; w.ldx = e->ldx; w.ldw = e->ldw;
  w.qrE = e->qrE; w.qrT = e->qrT; w.qrPerm = e->qrPerm; w.qr_tslot = e->qr_tslot;
  if (e->split_fused) {
    w.Xb = e->Xb;
    w.xb_slots = e->xb_slots;
  }
  return w;
}

// The Jacobi sweeps of one launch batch: nt tasks in workspace w, rows and
// columns bounded by maxR and maxC, launch shapes from tuning row tu, stop
// tolerance conv_tol.  *sweeps_run receives the sweeps checked, *launches
// the step kernels launched; no_conv(t) is called for each task the sweep
// limit stopped above the tolerance.  trace[] (gate, prepass, mode) labels
// the TNMPS_TRACE_SWEEPS lines.
template <typename NoConv>
tn_status jacobi_sweeps(tn_engine* e, const Workspace& w, int nt, size_t maxR, size_t maxC,
                        const tn_tune& tu, double conv_tol, const int (&trace)[3],
                        NoConv no_conv, int* sweeps_run, size_t* launches) {
  const int bw = tu.bw;
  const int nb = (int)((maxC + bw - 1) / bw);
  const int nbe = nb + (nb & 1);
  const int npairs = nbe / 2, steps = nbe - 1;
  const Tolerances& tol = e->tol;
  // One launch per round-robin step (two when the rows are split); the host
  // checks convergence per sweep.  Few pairs (a lone large SVD) leave most
  // CUs idle: split each pair's rows over S workgroups (gram kernel, then
  // solve + apply kernel).  The split takes two launches because every
  // workgroup's Gram matrix needs all of the pair's rows, which the other
  // workgroups of the pair overwrite in their apply: a kernel boundary
  // orders the reads before the writes.
  int S = 1;
  // (S > 1 needs nt * npairs <= cus, so nt <= cus: the partial Gram buffer
  // covers min(slots, cus) tasks)
  while (S < tu.split_max && (size_t)nt * npairs * S * 2 <= (size_t)2 * e->cus &&
         maxR >= (size_t)tu.split_rows * S && nt <= w.gpart_slots)
    S *= 2;
  JacobiArgs ja{};
  ja.X = w.X;
  ja.Xb = nullptr;
  ja.W = w.W;
  ja.gpart = w.gpart;
  ja.offs = w.offs;
  ja.info = w.info;
  // Many tasks: the step kernel takes the sweep's unconverged tasks from a
  // list that a one-workgroup kernel builds before the sweep, so tasks that
  // have converged (a lock-step batch runs until its slowest task has) cost
  // no workgroups.  The grid's rows are the unconverged count of the last
  // sweep read back (or nt), a bound on the list's length since a converged
  // task stays converged.  The arithmetic per task is the same either way.
  const bool use_list = S == 1 && nt >= e->list_min && nt <= 8192;
  ja.act = use_list ? w.act : nullptr;
  int open_known = nt;
  ja.xslot = (size_t)w.ldx * w.ldw;
  ja.wslot = (size_t)w.ldw * w.ldw;
  ja.gslot = w.gpart_per;
  ja.ldx = w.ldx;
  ja.ldw = w.ldw;
  ja.nt = nt;
  ja.nbe = nbe;
  ja.S = S;
  ja.Rh = (int)maxR;
  ja.Ch = (int)maxC;
  ja.tol_rot = tol.tol_rot;
  ja.inner_every = 1;
  ja.conv_tol = conv_tol;
  ja.tiny2 = tol.tiny_norm * tol.tiny_norm;
  // A row split of few tasks takes one launch per step when the workspace
  // has their second X buffers and the panels fit the kernel's registers.
  const bool fused = S > 1 && w.Xb && nt <= w.xb_slots && jacobi_fused_fits(ja, bw);
  if (fused) {
    ja.Xb = w.Xb;
    // One-launch row-split steps (a lone SVD or a few, panels of up to 256
    // rows: steps bound by the panel solve's latency) at the FP64 tolerance
    // rotate the pairs inside each 8-column block on every inner_every-th
    // step only.  Other launches keep every step: one workgroup per pair over
    // many tasks hides the solve behind other workgroups' traffic, and more
    // sweeps cost more there; the two-launch steps of 512-row panels lost
    // more to extra sweeps than they saved (FP64 N=48, chi=512: 1,020 -> 899
    // updates/s); and the looser FP32 / BF16-split tolerances end with a
    // residual near conv_tol^2 that the sparser schedule did not keep (it
    // moved the FP32 path's entropy error against FP64 over four N=64,
    // chi=256 layers past 1e-6).
    if (conv_tol <= tol.conv_tol) ja.inner_every = e->inner_every;
    // its rotation spreads 16-row tiles over the waves, so shorter slices
    // shorten it: split further while slices keep fused_rows rows and the
    // launch keeps one workgroup per CU (its LDS holds one per CU)
    while (S < tu.split_max && (size_t)nt * npairs * S * 2 <= (size_t)e->cus &&
           maxR >= (size_t)e->fused_rows * S * 2)
      S *= 2;
    ja.S = S;
  }
  // Sweeps are enqueued one ahead while the last sweep read back is still far
  // from convergence (max off-diagonal above kAheadOff: with quadratic
  // convergence the next sweep cannot reach conv_tol, so the one after it is
  // needed too); each sweep's maxima come back through pinned memory behind
  // an event.  Near convergence the host waits for every sweep, so no sweep
  // is enqueued that is not needed; a sweep enqueued after convergence
  // returns at once on the device.
  // (1e-3 for conv_tol = 1e-8; scaled by the square root for looser tolerances)
  const double kAheadOff = std::min(0.5, 1e-3 * std::sqrt(conv_tol / 1e-8));
  int launched = 0, checked = 0;
  size_t step_launches = 0;
  double last_worst = 1.0;
  bool done = false;
  const double* last_o = nullptr;  // the last sweep's maxima read back
  while (!done) {
    const int depth = last_worst > kAheadOff ? tu.ahead : 1;
    const auto h0 = std::chrono::steady_clock::now();
    while (launched < e->max_sweeps && launched - checked < depth) {
      if (use_list) launch_active(ja, launched, w.act, e->stream);
      for (int s = 0; s < steps; ++s) {
        if (S == 1) {
          launch_jacobi_step(ja, npairs, bw, launched, s, use_list ? open_known : nt,
                             e->stream);
          ++step_launches;
        } else if (fused) {
          launch_jacobi_fused(ja, npairs, launched, s, e->stream);
          ++step_launches;
        } else {
          launch_jacobi_gram(ja, npairs, bw, launched, s, e->stream);
          launch_jacobi_apply(ja, npairs, bw, launched, s, e->stream);
          step_launches += 2;
        }
      }
      TN_HIP(hipGetLastError());
      TN_HIP(hipMemcpyAsync(w.h_offs + (size_t)launched * nt, w.offs + (size_t)launched * nt,
                            sizeof(double) * nt, hipMemcpyDeviceToHost, e->stream));
      TN_HIP(hipEventRecord(e->sev[launched], e->stream));
      ++launched;
    }
    e->perf_host_launch_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    if (checked >= launched) break;  // sweep limit
    TN_HIP(hipEventSynchronize(e->sev[checked]));
    const double* o = w.h_offs + (size_t)checked * nt;
    last_o = o;
    done = true;
    double worst = 0.0;
    int open = 0;
    for (int t = 0; t < nt; ++t) {
      if (!(o[t] <= conv_tol)) {
        done = false;
        ++open;
      }
      worst = std::max(worst, o[t]);
    }
    if (e->trace_sweeps)
      std::fprintf(stderr,
                   "[tnmps] batch of %d (C <= %zu, gate %d, prepass %d, mode %d): sweep %d max "
                   "off-diagonal %.3e, %d not converged\n",
                   nt, maxC, trace[0], trace[1], trace[2], checked, worst, open);
    last_worst = worst;
    open_known = std::max(open, 1);
    ++checked;
  }
  // The sweep limit stopped tasks whose last sweep was still above the
  // tolerance: their results are written as they stand and reported.
  if (!done && last_o) {
    for (int t = 0; t < nt; ++t)
      if (!(last_o[t] <= conv_tol)) {
        no_conv(t);
        e->host_flags |= kFlagNoConv;
      }
  }
  // One-launch row split: a task's X is in Xb when it ran an odd number of
  // sweeps (a sweep is an odd number of steps, each writing the other
  // buffer).  It ran through the first sweep whose maxima were at or below
  // the tolerance (the device skips it after that), else every sweep
  // launched.
  if (fused) {
    for (int t = 0; t < nt; ++t) {
      int ran = launched;
      for (int j = 0; j < checked; ++j)
        if (w.h_offs[(size_t)j * nt + t] <= conv_tol) {
          ran = j + 1;
          break;
        }
      if (ran & 1)
        TN_HIP(hipMemcpyAsync(w.X + (size_t)t * ja.xslot, w.Xb + (size_t)t * ja.xslot,
                              sizeof(double2) * ja.xslot, hipMemcpyDeviceToDevice, e->stream));
    }
  }
  *sweeps_run = checked;
  *launches = step_launches;
  return TN_OK;
}

// The host's bounds on one task's SVD: columns C, rows R, and the outer
// bonds chiL, chiR.
struct TaskBounds {
  size_t C, R, chiL, chiR;
};

// Descriptor of update q in workspace slot t of w, and its bounds.
TaskBounds make_desc(const tn_engine* e, const Workspace& w, const Task& q, int t, UpdDesc& d) {
  tn_mps* h = q.h;
  const size_t k = q.k;
  std::memset(&d, 0, sizeof(d));
  d.mode = q.has_gate ? 0 : (q.pre ? q.pre : factor_mode(e, h, k));
  d.gk = h->gam + h->goff[k];
  d.gk1 = h->gam + h->goff[k + 1];
  d.lamL = k == 0 ? nullptr : h->lam + h->loff[k - 1];
  d.lamM = h->lam + h->loff[k];
  d.lamR = (k + 2 >= h->n) ? nullptr : h->lam + h->loff[k + 1];
  d.dims = h->dims + k;
  d.X = w.X + (size_t)t * w.ldx * w.ldw;
  d.W = w.W + (size_t)t * w.ldw * w.ldw;
  d.sraw = w.sraw + (size_t)t * w.ldw;
  d.sval = w.sval + (size_t)t * w.ldw;
  d.fac = w.fac + (size_t)t * w.ldw;
  d.perm = w.perm + (size_t)t * w.ldw;
  d.part = w.part + (size_t)t * 2 * kNPart;
  d.offs = w.offs + t;
  d.info = w.info + (size_t)t * kInfoLen;
  // the factored product goes to the second half of X's columns; the SVD
  // of a factored update uses C = K <= chi_max = ldw / 2 of them
  d.T = d.X + (size_t)(w.ldw / 2) * w.ldx;
  d.ldx = w.ldx;
  d.ldw = w.ldw;
  d.has_gate = q.has_gate ? 1 : 0;
  d.prec = h->prec;
  d.tkind = h->trunc.kind;
  d.tchi = (int)std::min<size_t>(h->trunc.chi, 1u << 30);
  d.cap_rank = (int)h->bmax[k];
  d.teps = h->trunc.eps;
  if (q.has_gate && !q.gdev)
    for (int i = 0; i < 16; ++i) d.gate[i] = make_double2(q.gate[2 * i], q.gate[2 * i + 1]);
  d.gsrc = q.has_gate ? q.gdev : nullptr;
  d.hflags = h->dflags();
  d.eflags = e->d_eflags;
  TaskBounds b;
  b.chiL = h->uL(k);
  b.chiR = h->uR(k + 1);
  // a factored request has C = K when K < min(m, n), else the full
  // C = min(m, n) <= K: either way C <= min(ub[k], m, n)
  b.C = std::min(2 * b.chiL, 2 * b.chiR);
  if (d.mode != 0) b.C = std::min(b.C, h->ub[k]);
  b.R = std::max(2 * b.chiL, 2 * b.chiR);
  return b;
}

// The host's view of a handle after update q: the bound on bond k and the
// freshness of sites k, k + 1.
void after_update(const Task& q) {
  tn_mps* h = q.h;
  const size_t k = q.k;
  const size_t chiL = h->uL(k), chiR = h->uR(k + 1);
  // a refactor's rank is at most min(K, m, n) <= ub[k]; a gate's min(m, n)
  const size_t b = std::min({2 * chiL, 2 * chiR, rank_bound(h), h->bmax[k]});
  h->ub[k] = q.has_gate ? b : std::min(b, h->ub[k]);
  // the prepass makes only its decomposed factor's site fresh
  h->fresh[k] = (q.pre == 1) ? 0 : kFreshLeft;
  h->fresh[k + 1] = (q.pre == 2) ? 0 : kFreshRight;
}

// QR preconditioner arguments for nt tasks of workspace w (task i in slot
// map[i], or slot i without a map).
QrArgs qr_args(const Workspace& w, int nt, const int* map) {
  QrArgs qa{};
  qa.X = w.X;
  qa.W = w.W;
  qa.E = w.qrE;
  qa.T = w.qrT;
  qa.cn = w.sraw;
  qa.perm0 = w.qrPerm;
  qa.info = w.info;
  qa.map = map;
  qa.xslot = (size_t)w.ldx * w.ldw;
  qa.wslot = (size_t)w.ldw * w.ldw;
  qa.tslot = w.qr_tslot;
  qa.ldx = w.ldx;
  qa.ldw = w.ldw;
  qa.nt = nt;
  return qa;
}

// Enqueue up to max_slots independent updates.  Within one handle, bonds
// must pairwise differ by at least 2 so that no update reads what another
// writes; different handles are independent.  With `ws` the batch runs in
// that workspace (one task, no allocation, no timing records) instead of the
// engine's.
tn_status run_chunk(tn_engine* e, const Task* tasks, int nt, const Workspace* ws = nullptr) {
  UpdDesc* hd = nullptr;
  int rs = 0;
  UpdDesc one{};
  if (ws) {
    if (nt != 1) return TN_INVALID_ARG;
    hd = &one;
  } else {
    TN_TRY(e->ensure_slots(nt));
    TN_TRY(e->upd.next(&hd, &rs));
  }
  const Workspace w = ws ? *ws : engine_workspace(e);
  size_t maxC = 1, maxR = 1, maxTiles = 1, maxCC = 1, maxChiL = 1, maxChiR = 1;
  bool any_factored = false;
  int mode0 = 0;  // first task's mode, for the sweep trace
  for (int t = 0; t < nt; ++t) {
    const TaskBounds tb = make_desc(e, w, tasks[t], t, hd[t]);
    if (t == 0) mode0 = hd[t].mode;
    any_factored = any_factored || hd[t].mode != 0;
    maxC = std::max(maxC, tb.C);
    maxR = std::max(maxR, tb.R);
    maxCC = std::max(maxCC, tb.C * tb.C);
    maxChiL = std::max(maxChiL, tb.chiL);
    maxChiR = std::max(maxChiR, tb.chiR);
  }
  const tn_tune& tu = e->tune_for(maxC);
  const size_t tv = (size_t)tu.theta_tile / 2;
  maxTiles = std::max<size_t>(1, ((maxChiL + tv - 1) / tv) * ((maxChiR + tv - 1) / tv));
  const UpdDesc* dd = nullptr;
  if (ws) {
    // a pageable source: the copy is staged before the call returns
    TN_HIP(hipMemcpyAsync(w.desc, hd, sizeof(UpdDesc), hipMemcpyHostToDevice, e->stream));
    dd = w.desc;
  } else {
    TN_TRY(e->upd.upload(rs, nt, e->stream));
    dd = e->upd.d[rs];
  }
  PerfRec* pr = nullptr;
  if (e->timing && !ws) {
    TN_TRY(e->perf_begin(0, nt, &pr));
    TN_HIP(hipEventRecord(pr->ev[0], e->stream));
  }
  // factored tasks fill X elementwise over the same grid
  const int gx_theta = (int)std::min<size_t>(std::max<size_t>(maxTiles, any_factored ? 256 : 1),
                                             1024);
  const int gx_w = (int)std::min<size_t>((maxCC + 255) / 256, 1024);
  launch_theta(dd, nt, tu.theta_tile, gx_theta, e->stream);
  if (pr) TN_HIP(hipEventRecord(pr->ev[8], e->stream));
  // QR preconditioner: the Jacobi sweeps then run on the C x C Y = R^H
  const bool use_qr = tu.qr && w.qrE && maxR <= (size_t)kQrMaxRows && maxC >= (size_t)e->qr_min;
  const QrArgs qa = qr_args(w, nt, nullptr);
  if (use_qr) {
    launch_qr(qa, (int)maxR, (int)maxC, e->stream);
    e->perf.qr_tasks += (size_t)nt;
  }
  if (pr) {
    TN_HIP(hipEventRecord(pr->ev[9], e->stream));
    pr->qr = use_qr;
  }
  launch_winit(dd, nt, gx_w, e->stream);
  TN_HIP(hipGetLastError());
  if (pr) TN_HIP(hipEventRecord(pr->ev[1], e->stream));
  // Block width 8 by default (measured faster than 16 even for a lone
  // 512-column SVD: the 32x32 inner solve costs more than the halved step
  // count saves); the tuning table may select the wide panels per bucket.
  // one precision per launch batch (run_tasks)
  const double conv_tol = e->tol.for_prec(tasks[0].h->prec);
  int sweeps_run = 0;
  size_t launches = 0;
  const int trace[3] = {tasks[0].has_gate ? 1 : 0, tasks[0].pre, mode0};
  if (pr) TN_HIP(hipEventRecord(pr->ev[2], e->stream));
  TN_TRY(jacobi_sweeps(e, w, nt, use_qr ? maxC : maxR, maxC, tu, conv_tol, trace,
                       [&](int t) { tasks[t].h->flags |= kFlagNoConv; }, &sweeps_run, &launches));
  if (pr) {
    TN_HIP(hipEventRecord(pr->ev[3], e->stream));
    pr->launches = launches;
  }
  if (use_qr) launch_qr_final(qa, (int)maxC, e->stream);
  if (pr) TN_HIP(hipEventRecord(pr->ev[4], e->stream));
  launch_finish(dd, nt, sweeps_run, e->stream);
  if (any_factored) launch_fmul(dd, nt, (int)maxR, (int)maxC, e->stream);
  if (pr) TN_HIP(hipEventRecord(pr->ev[6], e->stream));
  launch_split(dd, nt, e->stream);
  if (pr) TN_HIP(hipEventRecord(pr->ev[7], e->stream));
  launch_renorm(dd, nt, e->stream);
  TN_HIP(hipGetLastError());
  int n_updates = 0;
  for (int t = 0; t < nt; ++t) n_updates += tasks[t].pre ? 0 : 1;
  e->updates += n_updates;
  e->sweeps += (unsigned long long)sweeps_run;
  if (pr) {
    TN_HIP(hipEventRecord(pr->ev[5], e->stream));
    TN_HIP(hipMemcpyAsync(pr->h_info, w.info, sizeof(int) * kInfoLen * nt, hipMemcpyDeviceToHost,
                          e->stream));
    pr->gate.resize(nt);
    for (int t = 0; t < nt; ++t) pr->gate[t] = tasks[t].has_gate ? 1 : 0;
    pr->n_updates = n_updates;
    pr->sweeps = sweeps_run;
    pr->pending = true;
    TN_TRY(e->perf_drain(false));
  }
  for (int t = 0; t < nt; ++t) after_update(tasks[t]);
  return TN_OK;
}

// Tasks of different contraction precisions run as separate launch batches:
// each precision's SVDs stop at its own tolerance, so a task's sweeps do not
// depend on the batch it runs in.  (The tasks of one call are independent,
// so any grouping gives the same results.)
tn_status run_tasks(tn_engine* e, const std::vector<Task>& tasks) {
  bool mixed = false;
  for (const Task& t : tasks) mixed = mixed || t.h->prec != tasks[0].h->prec;
  if (mixed) {
    for (int prec : {TN_PREC_F64, TN_PREC_F32, TN_PREC_BF16X3}) {
      std::vector<Task> sub;
      for (const Task& t : tasks)
        if (t.h->prec == prec) sub.push_back(t);
      if (!sub.empty()) TN_TRY(run_tasks(e, sub));
    }
    return TN_OK;
  }
  const int nt = (int)tasks.size();
  for (int a = 0; a < nt; a += e->max_slots)
    TN_TRY(run_chunk(e, tasks.data() + a, std::min(e->max_slots, nt - a)));
  return TN_OK;
}

// Updates of one batch (bonds of one handle at least 2 apart), each refactor
// that needs one preceded by its gauge prepass.
tn_status run_updates(tn_engine* e, const std::vector<Task>& tasks) {
  std::vector<Task> pre;
  for (const Task& t : tasks) {
    if (t.has_gate) continue;
    const int m = prepass_mode(e, t.h, t.k);
    if (m) {
      Task p = t;
      p.pre = m;
      pre.push_back(p);
    }
  }
  if (!pre.empty()) TN_TRY(run_tasks(e, pre));
  return run_tasks(e, tasks);
}

tn_status refactor(tn_mps* h, size_t k) {
  Task t;
  t.h = h;
  t.k = k;
  t.has_gate = false;
  return run_updates(h->eng, {t});
}

tn_status run_measures(tn_engine* e, const std::vector<MTask>& ms) {
  const int nm = (int)ms.size();
  if (nm == 0) return TN_OK;
  TN_TRY(e->meas.ensure(nm, e->stream));
  MeasDesc* hd = nullptr;
  int rs = 0;
  TN_TRY(e->meas.next(&hd, &rs));
  for (int i = 0; i < nm; ++i) {
    const MTask& q = ms[i];
    tn_mps* h = q.h;
    MeasDesc& d = hd[i];
    std::memset(&d, 0, sizeof(d));
    d.g = h->gam + h->goff[q.k];
    d.lamL = q.k == 0 ? nullptr : h->lam + h->loff[q.k - 1];
    d.lamR = (q.k + 1 >= h->n) ? nullptr : h->lam + h->loff[q.k];
    d.dims = h->dims + q.k;
    d.forced = q.forced;
    if (q.forced != -2) h->fresh[q.k] = 0;  // projected
    d.slot = q.slot;
    d.r = q.r;
    d.out_outcome = e->d_out;
    d.out_prob = e->d_prob;
    d.out_probs = e->d_probs;
    d.out_flag = e->d_flag;
  }
  if (e->timing) TN_TRY(e->ensure_mdims(nm));
  for (int i = 0; i < nm && e->timing; ++i) hd[i].out_dims = e->d_mdims;
  TN_TRY(e->meas.upload(rs, nm, e->stream));
  PerfRec* pr = nullptr;
  if (e->timing) {
    TN_TRY(e->perf_begin(1, nm, &pr));
    pr->proj.resize(nm);
    for (int i = 0; i < nm; ++i) pr->proj[i] = ms[i].forced != -2 ? 1 : 0;
    TN_HIP(hipEventRecord(pr->ev[0], e->stream));
  }
  // workgroups per site: at least 8 rows each, from the host's bound on chiL
  size_t maxL = 1;
  for (const MTask& q : ms) maxL = std::max(maxL, q.h->uL(q.k));
  const int G = (int)std::min<size_t>(kMeasParts, std::max<size_t>(1, maxL / 8));
  launch_measure(e->meas.d[rs], nm, G, e->d_mpart, e->stream);
  TN_HIP(hipGetLastError());
  if (pr) {
    TN_HIP(hipEventRecord(pr->ev[1], e->stream));
    TN_HIP(hipMemcpyAsync(pr->h_info, e->d_mdims, sizeof(int) * 2 * nm, hipMemcpyDeviceToHost,
                          e->stream));
    TN_HIP(hipEventRecord(pr->ev[2], e->stream));  // the dims have arrived
    pr->pending = true;
    TN_TRY(e->perf_drain(false));
  }
  return TN_OK;
}

// All read-outs in one launch, results to d_scalar[slot].
tn_status run_entropies(tn_engine* e, const std::vector<EntReq>& es) {
  if (es.empty()) return TN_OK;
  const int ne = (int)es.size();
  TN_TRY(e->ents.ensure(ne, e->stream));
  EntDesc* hd = nullptr;
  int rs = 0;
  TN_TRY(e->ents.next(&hd, &rs));
  std::copy(es.begin(), es.end(), hd);
  TN_TRY(e->ents.upload(rs, ne, e->stream));
  launch_entropy(e->ents.d[rs], ne, e->d_scalar, e->stream);
  TN_HIP(hipGetLastError());
  return TN_OK;
}

tn_status refactor_sweep(tn_mps* h) {
  // mps.rs:894-897
  if (h->n < 2) return TN_OK;
  for (size_t k = 0; k + 1 < h->n; ++k) TN_TRY(refactor(h, k));
  for (size_t k = h->n - 2; k-- > 0;) TN_TRY(refactor(h, k));
  return TN_OK;
}

// Enqueue the copy of the engine's device flag word into pinned memory; its
// value is valid after the next synchronisation of the stream.
tn_status fetch_eflags(tn_engine* e) {
  TN_HIP(hipMemcpyAsync(e->h_eflags, e->d_eflags, sizeof(int), hipMemcpyDeviceToHost, e->stream));
  return TN_OK;
}

tn_status read_results(tn_engine* e, int base, int count, int* outcomes, double* probs,
                       double* both, int* flags) {
  if (count > 0) {
    if (outcomes)
      TN_HIP(hipMemcpyAsync(outcomes, e->d_out + base, sizeof(int) * count,
                            hipMemcpyDeviceToHost, e->stream));
    if (probs)
      TN_HIP(hipMemcpyAsync(probs, e->d_prob + base, sizeof(double) * count,
                            hipMemcpyDeviceToHost, e->stream));
    if (both)
      TN_HIP(hipMemcpyAsync(both, e->d_probs + 2 * base, sizeof(double) * 2 * count,
                            hipMemcpyDeviceToHost, e->stream));
    if (flags)
      TN_HIP(hipMemcpyAsync(flags, e->d_flag + base, sizeof(int) * count,
                            hipMemcpyDeviceToHost, e->stream));
  }
  TN_TRY(fetch_eflags(e));
  TN_HIP(hipStreamSynchronize(e->stream));
  return TN_OK;
}

// ------------------------------------------------------------ schedules
// One step of a handle's schedule: a refactor of bond k, or a measurement.
struct Op {
  bool meas;
  size_t k;
  int forced;
  double r;
  int slot;
};

// The schedule of apply_meas_prob_layer (mps.rs:2206-2252): drop out-of-range
// sites (:2212-2216), stable sort by site (:2218), dedup keeping the first
// (:2219), refactor last..k before each measurement (:2225-2228), then the
// trailing L->R range last..n-1 (:2246-2249) and the R->L sweep over
// (0..n-2).rev() (:2250).  Result slots are assigned from `slot` upward.
void meas_schedule(const tn_mps* h, size_t m, const size_t* ks, const int* forced,
                   const double* r, int* slot, std::vector<Op>* ops, std::vector<size_t>* done) {
  struct M {
    size_t k;
    int f;
    double r;
  };
  std::vector<M> ms;
  for (size_t i = 0; i < m; ++i) {
    if (ks[i] >= h->n) continue;
    ms.push_back({ks[i], forced ? forced[i] : -1, r ? r[i] : 0.0});
  }
  if (ms.empty()) return;
  std::stable_sort(ms.begin(), ms.end(), [](const M& a, const M& b) { return a.k < b.k; });
  std::vector<M> dd;
  for (const M& x : ms)
    if (dd.empty() || dd.back().k != x.k) dd.push_back(x);
  bool have_last = false;
  size_t last = 0;
  for (const M& x : dd) {
    if (have_last)
      for (size_t j = last; j < x.k; ++j) ops->push_back({false, j, 0, 0.0, 0});
    ops->push_back({true, x.k, x.f >= 0 ? x.f : -1, x.r, (*slot)++});
    done->push_back(x.k);
    have_last = true;
    last = x.k;
  }
  if (have_last)
    for (size_t j = last; j + 1 < h->n; ++j) ops->push_back({false, j, 0, 0.0, 0});
  if (h->n >= 2)
    for (size_t j = h->n - 2; j-- > 0;) ops->push_back({false, j, 0, 0.0, 0});
}

// Gate groups of one handle: consecutive two-qubit gates on bonds at least 2
// apart form one group (they touch disjoint sites and bonds); a one-qubit
// gate is a group of its own.  Out-of-range gates are dropped (mps.rs:1301,
// 1335).
struct GateOp {
  size_t k;
  int dim;
  const double* op;
};

tn_status gate_groups(const tn_mps* h, size_t m, const size_t* ks, const int* dimsz,
                      const double* ops, std::vector<std::vector<GateOp>>* groups) {
  size_t off = 0;
  std::vector<GateOp> cur;
  for (size_t i = 0; i < m; ++i) {
    const int dsz = dimsz ? dimsz[i] : 4;
    const double* op = ops + off;
    if (dsz == 2) {
      off += 8;
      if (ks[i] >= h->n) continue;
      if (!cur.empty()) groups->push_back(cur), cur.clear();
      groups->push_back({{ks[i], 2, op}});
    } else if (dsz == 4) {
      off += 32;
      if (h->n == 1 || ks[i] >= h->n - 1) continue;
      bool conflict = false;
      for (const GateOp& b : cur)
        if ((b.k > ks[i] ? b.k - ks[i] : ks[i] - b.k) < 2) conflict = true;
      if (conflict) groups->push_back(cur), cur.clear();
      cur.push_back({ks[i], 4, op});
    } else {
      return TN_OPERATOR_INCOMPATIBLE_SHAPE;  // mps.rs:2104
    }
  }
  if (!cur.empty()) groups->push_back(cur);
  return TN_OK;
}

// Advance several handles of one engine through their gate groups in lock
// step: group j of every handle runs in one launch batch.
tn_status exec_gates(tn_engine* e, const std::vector<tn_mps*>& hs,
                     const std::vector<std::vector<std::vector<GateOp>>>& groups) {
  size_t depth = 0;
  for (const auto& g : groups) depth = std::max(depth, g.size());
  for (size_t j = 0; j < depth; ++j) {
    std::vector<Task> tasks;
    for (size_t i = 0; i < hs.size(); ++i) {
      if (j >= groups[i].size()) continue;
      for (const GateOp& g : groups[i][j]) {
        if (g.dim == 2) {
          launch_u1(hs[i]->gam + hs[i]->goff[g.k], hs[i]->dims + g.k, g.op, 64, e->stream);
          hs[i]->fresh[g.k] = 0;  // the operator need not be unitary
          TN_HIP(hipGetLastError());
        } else {
          Task t;
          t.h = hs[i];
          t.k = g.k;
          t.has_gate = true;
          std::memcpy(t.gate, g.op, sizeof(double) * 32);
          tasks.push_back(t);
        }
      }
    }
    TN_TRY(run_updates(e, tasks));
  }
  return TN_OK;
}

// ---- SVD task pool ---------------------------------------------------------
// The measurement schedules of many handles as a pool of SVD tasks instead
// of lock-step batches.  A handle has at most one refactor in flight (its
// schedule is sequential); every pool iteration runs one Jacobi sweep over
// all tasks in flight; a task whose sweep stayed at or below its tolerance
// is finished at once, and its handle's next operations (a measurement, the
// next refactor or that refactor's gauge prepass) start in the freed slot on
// the next iteration.  A lock-step batch runs until its slowest task has
// converged, so its last sweeps hold a few tasks; the pool keeps its slots
// filled.  Each task's arithmetic is the lock-step path's but for the
// round-robin block count (from the widest task in flight) and the launch
// shape (one workgroup per block pair), so results agree with solo runs to
// rounding.  The per-sweep maxima live in row 0 of the workspace's offs
// (slot t at offs[t]), cleared before every sweep.
struct PoolSlot {
  Task task{};
  int handle = -1;   // index into the pool's handles; -1: free
  int sweeps = 0;
  bool qr = false;
  double tol = 0.0;
  TaskBounds b{};
  UpdDesc desc{};
};

// Contraction (or factored fill), QR preconditioner and W = I for newly
// admitted tasks (slots adm).
tn_status pool_admit(tn_engine* e, const Workspace& w, std::vector<PoolSlot>& sl,
                     const std::vector<int>& adm) {
  const int n = (int)adm.size();
  UpdDesc* hd = nullptr;
  int rs = 0;
  TN_TRY(e->upd.next(&hd, &rs));
  size_t maxC = 1, maxR = 1, maxCC = 1, maxChiL = 1, maxChiR = 1;
  bool any_factored = false;
  for (int j = 0; j < n; ++j) {
    PoolSlot& p = sl[adm[j]];
    p.b = make_desc(e, w, p.task, adm[j], hd[j]);
    p.desc = hd[j];
    p.tol = e->tol.for_prec(p.task.h->prec);
    any_factored = any_factored || hd[j].mode != 0;
    maxC = std::max(maxC, p.b.C);
    maxR = std::max(maxR, p.b.R);
    maxCC = std::max(maxCC, p.b.C * p.b.C);
    maxChiL = std::max(maxChiL, p.b.chiL);
    maxChiR = std::max(maxChiR, p.b.chiR);
    e->h_pmap[0][j] = adm[j];
  }
  TN_TRY(e->upd.upload(rs, n, e->stream));
  const UpdDesc* dd = e->upd.d[rs];
  const tn_tune& tu = e->tune_for(maxC);
  const bool use_qr = tu.qr && maxR <= (size_t)kQrMaxRows && maxC >= (size_t)e->qr_min;
  for (int s : adm) sl[s].qr = use_qr;
  PerfRec* pr = nullptr;
  if (e->timing) {
    TN_TRY(e->perf_begin(0, 0, &pr));
    TN_HIP(hipEventRecord(pr->ev[0], e->stream));
  }
  const size_t tv = (size_t)tu.theta_tile / 2;
  const size_t maxTiles =
      std::max<size_t>(1, ((maxChiL + tv - 1) / tv) * ((maxChiR + tv - 1) / tv));
  const int gx_theta = (int)std::min<size_t>(std::max<size_t>(maxTiles, any_factored ? 256 : 1),
                                             1024);
  const int gx_w = (int)std::min<size_t>((maxCC + 255) / 256, 1024);
  launch_theta(dd, n, tu.theta_tile, gx_theta, e->stream);
  if (pr) TN_HIP(hipEventRecord(pr->ev[8], e->stream));
  if (use_qr) {
    TN_HIP(hipMemcpyAsync(e->d_pmap[0], e->h_pmap[0], sizeof(int) * n, hipMemcpyHostToDevice,
                          e->stream));
    launch_qr(qr_args(w, n, e->d_pmap[0]), (int)maxR, (int)maxC, e->stream);
    e->perf.qr_tasks += (size_t)n;
  }
  if (pr) TN_HIP(hipEventRecord(pr->ev[9], e->stream));
  launch_winit(dd, n, gx_w, e->stream);
  TN_HIP(hipGetLastError());
  if (pr) {
    for (int i : {1, 2, 3, 4, 6, 7, 5}) TN_HIP(hipEventRecord(pr->ev[i], e->stream));
    pr->pending = true;
    TN_TRY(e->perf_drain(false));
  }
  return TN_OK;
}

// One Jacobi sweep over the tasks in slots act (ascending); their maxima
// come back in w.h_offs[slot].
tn_status pool_sweep(tn_engine* e, const Workspace& w, const std::vector<PoolSlot>& sl,
                     const std::vector<int>& act, int cap) {
  const int n = (int)act.size();
  size_t maxC = 1, maxRJ = 1;
  e->h_pact[0] = n;
  for (int j = 0; j < n; ++j) {
    const PoolSlot& p = sl[act[j]];
    e->h_pact[1 + j] = act[j];
    maxC = std::max(maxC, p.b.C);
    maxRJ = std::max(maxRJ, p.qr ? p.b.C : p.b.R);
  }
  TN_HIP(hipMemcpyAsync(w.act, e->h_pact, sizeof(int) * (n + 1), hipMemcpyHostToDevice,
                        e->stream));
  TN_HIP(hipMemsetAsync(w.offs, 0, sizeof(double) * cap, e->stream));
  const int nb = (int)((maxC + kBW - 1) / kBW);
  const int nbe = nb + (nb & 1);
  JacobiArgs ja{};
  ja.X = w.X;
  ja.W = w.W;
  ja.gpart = w.gpart;
  ja.offs = w.offs;
  ja.info = w.info;
  ja.act = w.act;
  ja.xslot = (size_t)w.ldx * w.ldw;
  ja.wslot = (size_t)w.ldw * w.ldw;
  ja.gslot = w.gpart_per;
  ja.ldx = w.ldx;
  ja.ldw = w.ldw;
  ja.nt = cap;
  ja.nbe = nbe;
  ja.S = 1;
  ja.Rh = (int)maxRJ;
  ja.Ch = (int)maxC;
  ja.tol_rot = e->tol.tol_rot;
  ja.inner_every = 1;  // one workgroup per pair (see jacobi_sweeps)
  ja.conv_tol = 0.0;  // (sweep index 0: no task counts as converged in the kernels)
  ja.tiny2 = e->tol.tiny_norm * e->tol.tiny_norm;
  PerfRec* pr = nullptr;
  if (e->timing) {
    TN_TRY(e->perf_begin(2, 0, &pr));
    TN_HIP(hipEventRecord(pr->ev[0], e->stream));
  }
  const auto h0 = std::chrono::steady_clock::now();
  for (int s = 0; s < nbe - 1; ++s) launch_jacobi_step(ja, nbe / 2, kBW, 0, s, n, e->stream);
  TN_HIP(hipGetLastError());
  e->perf_host_launch_ms +=
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
  if (pr) {
    TN_HIP(hipEventRecord(pr->ev[1], e->stream));
    pr->launches = (size_t)(nbe - 1);
    pr->pending = true;
  }
  TN_HIP(hipMemcpyAsync(w.h_offs, w.offs, sizeof(double) * cap, hipMemcpyDeviceToHost,
                        e->stream));
  TN_HIP(hipEventRecord(e->sev[0], e->stream));
  TN_HIP(hipEventSynchronize(e->sev[0]));
  e->sweeps += 1;
  if (pr) TN_TRY(e->perf_drain(false));
  return TN_OK;
}

// Singular values, rank rule, the QR's final step, factor products, split
// and renormalisation for the finished tasks in slots ret.
tn_status pool_retire(tn_engine* e, const Workspace& w, std::vector<PoolSlot>& sl,
                      const std::vector<int>& ret, int cap) {
  const int n = (int)ret.size();
  UpdDesc* hd = nullptr;
  int rs = 0;
  TN_TRY(e->upd.next(&hd, &rs));
  size_t maxC = 1, maxR = 1;
  int nq = 0, maxsw = 0, n_updates = 0;
  bool any_factored = false;
  for (int j = 0; j < n; ++j) {
    const PoolSlot& p = sl[ret[j]];
    hd[j] = p.desc;
    maxC = std::max(maxC, p.b.C);
    maxR = std::max(maxR, p.b.R);
    maxsw = std::max(maxsw, p.sweeps);
    any_factored = any_factored || p.desc.mode != 0;
    n_updates += p.task.pre ? 0 : 1;
    if (p.qr) e->h_pmap[1][nq++] = ret[j];
  }
  TN_TRY(e->upd.upload(rs, n, e->stream));
  const UpdDesc* dd = e->upd.d[rs];
  PerfRec* pr = nullptr;
  if (e->timing) {
    TN_TRY(e->perf_begin(0, cap, &pr));
    for (int i : {0, 8, 9, 1, 2, 3}) TN_HIP(hipEventRecord(pr->ev[i], e->stream));
  }
  if (nq) {
    TN_HIP(hipMemcpyAsync(e->d_pmap[1], e->h_pmap[1], sizeof(int) * nq, hipMemcpyHostToDevice,
                          e->stream));
    launch_qr_final(qr_args(w, nq, e->d_pmap[1]), (int)maxC, e->stream);
  }
  if (pr) TN_HIP(hipEventRecord(pr->ev[4], e->stream));
  launch_finish(dd, n, maxsw, e->stream);
  if (any_factored) launch_fmul(dd, n, (int)maxR, (int)maxC, e->stream);
  if (pr) TN_HIP(hipEventRecord(pr->ev[6], e->stream));
  launch_split(dd, n, e->stream);
  if (pr) TN_HIP(hipEventRecord(pr->ev[7], e->stream));
  launch_renorm(dd, n, e->stream);
  TN_HIP(hipGetLastError());
  e->updates += n_updates;
  if (pr) {
    TN_HIP(hipEventRecord(pr->ev[5], e->stream));
    TN_HIP(hipMemcpyAsync(pr->h_info, w.info, sizeof(int) * kInfoLen * cap,
                          hipMemcpyDeviceToHost, e->stream));
    pr->nt = n;
    pr->slots = ret;
    pr->gate.assign(n, 0);
    pr->n_updates = n_updates;
    pr->pending = true;
    TN_TRY(e->perf_drain(false));
  }
  for (int s : ret) after_update(sl[s].task);
  return TN_OK;
}

tn_status run_pool(tn_engine* e, const std::vector<tn_mps*>& hs,
                   const std::vector<std::vector<Op>>& ops) {
  const int H = (int)hs.size();
  const int cap = std::min(H, e->max_slots);
  TN_TRY(e->ensure_slots(cap));
  TN_TRY(e->ensure_pool(cap));
  const Workspace w = engine_workspace(e);
  std::vector<size_t> pos(H, 0);
  std::vector<int> busy(H, -1);
  std::vector<char> pre_done(H, 0);  // the gauge prepass of the refactor at pos has run
  std::vector<PoolSlot> sl(cap);
  std::vector<int> freeq;
  for (int s = cap - 1; s >= 0; --s) freeq.push_back(s);
  int inflight = 0;
  for (;;) {
    // idle handles: a measurement due next runs first (it clears the
    // freshness its site had, which the next refactor's mode reads) ...
    std::vector<MTask> ms;
    for (int i = 0; i < H; ++i) {
      if (busy[i] >= 0 || pos[i] >= ops[i].size() || !ops[i][pos[i]].meas) continue;
      const Op& o = ops[i][pos[i]];
      ms.push_back({hs[i], o.k, o.forced, o.r, o.slot});
      ++pos[i];
    }
    TN_TRY(run_measures(e, ms));
    // ... then a refactor due next (or its gauge prepass) takes a free slot
    std::vector<int> adm;
    bool pending_ops = false;
    for (int i = 0; i < H; ++i) {
      if (busy[i] >= 0 || pos[i] >= ops[i].size()) continue;
      pending_ops = true;
      if (ops[i][pos[i]].meas || freeq.empty()) continue;  // next iteration
      Task t{};
      t.h = hs[i];
      t.k = ops[i][pos[i]].k;
      t.has_gate = false;
      t.pre = pre_done[i] ? 0 : prepass_mode(e, hs[i], t.k);
      const int s = freeq.back();
      freeq.pop_back();
      sl[s].task = t;
      sl[s].handle = i;
      sl[s].sweeps = 0;
      busy[i] = s;
      adm.push_back(s);
    }
    if (!adm.empty()) TN_TRY(pool_admit(e, w, sl, adm));
    inflight += (int)adm.size();
    if (inflight == 0) {
      if (!pending_ops) break;
      continue;
    }
    std::vector<int> act;
    for (int s = 0; s < cap; ++s)
      if (sl[s].handle >= 0) act.push_back(s);
    TN_TRY(pool_sweep(e, w, sl, act, cap));
    std::vector<int> ret;
    for (int s : act) {
      PoolSlot& p = sl[s];
      ++p.sweeps;
      if (!(w.h_offs[s] <= p.tol)) {
        if (p.sweeps < e->max_sweeps) continue;
        // the sweep limit: written as it stands and reported
        p.task.h->flags |= kFlagNoConv;
        e->host_flags |= kFlagNoConv;
      }
      ret.push_back(s);
    }
    if (!ret.empty()) TN_TRY(pool_retire(e, w, sl, ret, cap));
    for (int s : ret) {
      const int i = sl[s].handle;
      if (sl[s].task.pre) {
        pre_done[i] = 1;
      } else {
        pre_done[i] = 0;
        ++pos[i];
      }
      busy[i] = -1;
      sl[s].handle = -1;
      freeq.push_back(s);
      --inflight;
    }
  }
  return TN_OK;
}

// Advance several handles through their measurement schedules: as an SVD
// task pool from pool_min handles on (when pool_min is set), else in lock
// step.
tn_status exec_ops(tn_engine* e, const std::vector<tn_mps*>& hs,
                   const std::vector<std::vector<Op>>& ops) {
  if (e->pool_min >= 2 && (int)hs.size() >= e->pool_min) return run_pool(e, hs, ops);
  size_t depth = 0;
  for (const auto& o : ops) depth = std::max(depth, o.size());
  for (size_t j = 0; j < depth; ++j) {
    std::vector<Task> tasks;
    std::vector<MTask> ms;
    for (size_t i = 0; i < hs.size(); ++i) {
      if (j >= ops[i].size()) continue;
      const Op& o = ops[i][j];
      if (o.meas) {
        ms.push_back({hs[i], o.k, o.forced, o.r, o.slot});
      } else {
        Task t;
        t.h = hs[i];
        t.k = o.k;
        t.has_gate = false;
        tasks.push_back(t);
      }
    }
    TN_TRY(run_measures(e, ms));
    TN_TRY(run_updates(e, tasks));
  }
  return TN_OK;
}

tn_status check_flags(const std::vector<int>& flags) {
  for (int f : flags)
    if (f & 2) return TN_SAMPLE_RANGE;
  for (int f : flags)
    if (f & 1) return TN_ZERO_PROB;
  return TN_OK;
}

// Allocation bound per bond of a new handle: chi_cap when given, chi for
// Const(chi), else the largest exact bond dimension min(2^{b+1}, 2^{n-b-1})
// up to kMaxCols / 2 (the widest SVD the library poses).  A None or Cutoff
// state whose rank rule asks for more than its bound is truncated to it and
// the call reports TN_CAPACITY.
size_t default_cap(size_t n, tn_trunc trunc, size_t chi_cap) {
  if (chi_cap) return chi_cap;
  if (trunc.kind == TN_TRUNC_CONST) return std::max<size_t>(trunc.chi, 1);
  size_t m = 1;
  for (size_t b = 0; b + 1 < n; ++b) m = std::max(m, exact_bond_max(n, b));
  return std::min<size_t>(m, (size_t)kMaxCols / 2);
}

// After a synchronisation that followed fetch_eflags: the failure flags of
// the updates enqueued since the last collection.  Device flags (a clamped
// rank) are copied to the handles of `hs` that set them, which keep them
// until tn_take_flags; the device words are cleared.  Returns TN_CAPACITY for
// a clamped rank, else TN_NO_CONVERGENCE for an SVD stopped at the sweep
// limit, else TN_OK.
tn_status collect_flags(tn_engine* e, tn_mps* const* hs, size_t count) {
  int f = *e->h_eflags | e->host_flags;
  e->host_flags = 0;
  if (*e->h_eflags) {
    static const int zero = 0;
    for (size_t i = 0; i < count; ++i) {
      tn_mps* h = hs[i];
      int hf = 0;
      TN_HIP(hipMemcpy(&hf, h->dflags(), sizeof(int), hipMemcpyDeviceToHost));
      if (hf) {
        h->flags |= hf;
        TN_HIP(hipMemcpy(h->dflags(), &zero, sizeof(int), hipMemcpyHostToDevice));
      }
    }
    *e->h_eflags = 0;
    TN_HIP(hipMemcpy(e->d_eflags, &zero, sizeof(int), hipMemcpyHostToDevice));
  }
  if (f & kFlagCapacity) return TN_CAPACITY;
  if (f & kFlagNoConv) return TN_NO_CONVERGENCE;
  return TN_OK;
}
tn_status collect_flags(tn_engine* e, const std::vector<tn_mps*>& hs) {
  return collect_flags(e, hs.data(), hs.size());
}

// Synchronise the engine's stream and collect the failure flags (for calls
// that return no other results).
tn_status sync_collect(tn_engine* e, const std::vector<tn_mps*>& hs) {
  TN_TRY(fetch_eflags(e));
  TN_HIP(hipStreamSynchronize(e->stream));
  return collect_flags(e, hs);
}

// The status of a call: measurement statuses first, then failure flags.
tn_status first_status(tn_status a, tn_status b) { return a != TN_OK ? a : b; }

tn_status new_handle(tn_engine* e, size_t n, tn_trunc trunc, size_t chi_cap, tn_mps** out) {
  const size_t cap = default_cap(n, trunc, chi_cap);
  tn_mps* h = new tn_mps();
  h->n = n;
  h->trunc = trunc;
  h->cap = cap;
  h->bmax.resize(n > 1 ? n - 1 : 0);
  size_t maxb = 1;
  for (size_t b = 0; b + 1 < n; ++b) {
    h->bmax[b] = std::min(cap, exact_bond_max(n, b));
    maxb = std::max(maxb, h->bmax[b]);
  }
  if (maxb > e->chi_max) {
    g_last_error = "bond allocation above the engine's chi_max";
    delete h;
    return TN_INVALID_ARG;
  }
  h->goff.resize(n);
  size_t g = 0;
  for (size_t k = 0; k < n; ++k) {
    h->goff[k] = g;
    g += h->bL(k) * 2 * h->bR(k);
  }
  h->gam_elems = g;
  h->loff.resize(n > 1 ? n - 1 : 0);
  size_t l = 0;
  for (size_t b = 0; b + 1 < n; ++b) {
    h->loff[b] = l;
    l += h->bmax[b];
  }
  h->lam_elems = std::max<size_t>(l, 1);
  h->ub.assign(n > 1 ? n - 1 : 0, 1);
  h->fresh.assign(n, kFreshLeft | kFreshRight);
  h->eng = e;
  e->users++;
  auto fail = [&](hipError_t er) {
    g_last_error = hipGetErrorString(er);
    delete h;
    return TN_DEVICE;
  };
  hipError_t er;
  if ((er = hipMalloc(&h->gam, sizeof(double2) * h->gam_elems)) != hipSuccess) return fail(er);
  if ((er = hipMalloc(&h->lam, sizeof(double) * h->lam_elems)) != hipSuccess) return fail(er);
  // n + 1 bond dimensions, then the handle's sticky flag word
  if ((er = hipMalloc(&h->dims, sizeof(int) * (n + 2))) != hipSuccess) return fail(er);
  // |0...0>: Gamma[k] = (1,2,1) with [0,0,0] = 1, lambda = [1] (mps.rs:378-388)
  std::vector<double2> g0(h->gam_elems, make_double2(0.0, 0.0));
  for (size_t k = 0; k < n; ++k) g0[h->goff[k]] = make_double2(1.0, 0.0);
  std::vector<double> l0(h->lam_elems, 0.0);
  for (size_t b = 0; b + 1 < n; ++b) l0[h->loff[b]] = 1.0;
  std::vector<int> d0(n + 2, 1);
  d0[n + 1] = 0;
  if ((er = hipMemcpyAsync(h->gam, g0.data(), sizeof(double2) * h->gam_elems,
                           hipMemcpyHostToDevice, e->stream)) != hipSuccess)
    return fail(er);
  if ((er = hipMemcpyAsync(h->lam, l0.data(), sizeof(double) * h->lam_elems,
                           hipMemcpyHostToDevice, e->stream)) != hipSuccess)
    return fail(er);
  if ((er = hipMemcpyAsync(h->dims, d0.data(), sizeof(int) * (n + 2), hipMemcpyHostToDevice,
                           e->stream)) != hipSuccess)
    return fail(er);
  if ((er = hipStreamSynchronize(e->stream)) != hipSuccess) return fail(er);
  *out = h;
  return TN_OK;
}

// Run the measurement schedules of a batch in lock step, then the
// entropies, and read back outcomes, probabilities and counts.
tn_status finish_layers(tn_engine* e, const std::vector<tn_mps*>& hv,
                        const std::vector<std::vector<Op>>& mops,
                        const std::vector<size_t>& counts, int slot, const size_t* ent_bonds,
                        double* ent_out, size_t* out_counts, int* out_outcomes,
                        double* out_probs) {
  const size_t count = hv.size();
  TN_TRY(exec_ops(e, hv, mops));
  std::vector<EntReq> es;
  if (ent_out && ent_bonds) {
    for (size_t i = 0; i < count; ++i) {
      tn_mps* h = hv[i];
      if (h->n >= 2 && ent_bonds[i] < h->n - 1)
        es.push_back({h->lam + h->loff[ent_bonds[i]], h->dims + ent_bonds[i] + 1, 1.0, (int)i});
    }
    TN_TRY(run_entropies(e, es));
  }
  std::vector<int> flags(slot);
  TN_TRY(read_results(e, 0, slot, out_outcomes, out_probs, nullptr, flags.data()));
  if (out_counts)
    for (size_t i = 0; i < count; ++i) out_counts[i] = counts[i];
  if (ent_out && ent_bonds) {
    std::vector<double> tmp(count, 0.0);
    TN_HIP(hipMemcpy(tmp.data(), e->d_scalar, sizeof(double) * count, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < count; ++i)
      ent_out[i] = (hv[i]->n >= 2 && ent_bonds[i] < hv[i]->n - 1) ? tmp[i] : NAN;
  }
  return first_status(check_flags(flags), collect_flags(e, hv));
}

tn_status check_trunc(tn_trunc t) {
  return (t.kind < 0 || t.kind > 2) ? TN_INVALID_ARG : TN_OK;
}

tn_status check_meas_args(size_t m, const int* forced, const double* r) {
  for (size_t i = 0; i < m; ++i) {
    if (forced && (forced[i] < -1 || forced[i] > 1)) return TN_INVALID_ARG;
    if ((!forced || forced[i] < 0) && !r) return TN_INVALID_ARG;
  }
  return TN_OK;
}

}  // namespace

// ==========================================================================
// C ABI
// ==========================================================================
extern "C" {

const char* tn_status_str(tn_status s) {
  switch (s) {
    case TN_OK: return "ok";
    case TN_EMPTY_SYSTEM: return "empty system";
    case TN_INVALID_QUANTUM_NUMBER: return "invalid quantum number";
    case TN_UNPHYSICAL_INDEX: return "unphysical index";
    case TN_STATE_INCOMPATIBLE_SHAPE: return "state incompatible shape";
    case TN_OPERATOR_INCOMPATIBLE_SHAPE: return "operator incompatible shape";
    case TN_NONE: return "none (index out of range)";
    case TN_ZERO_PROB: return "post-selected outcome has zero probability";
    case TN_SAMPLE_RANGE: return "uniform draw at or above the total probability";
    case TN_CAPACITY: return "rank exceeded the bond allocation";
    case TN_DEVICE: return "HIP runtime error";
    case TN_INVALID_ARG: return "invalid argument";
    case TN_NO_CONVERGENCE: return "SVD did not converge";
  }
  return "unknown status";
}

const char* tn_last_error(void) { return g_last_error.c_str(); }

int tn_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

tn_status tn_engine_new(size_t chi_max, void* stream, tn_engine** out) {
  if (!out) return TN_INVALID_ARG;
  *out = nullptr;
  if (chi_max == 0 || 2 * chi_max > (size_t)kMaxCols) {
    // the singular-value sort of k_finish holds at most kMaxCols values
    g_last_error = "chi_max must be in 1..512";
    return TN_INVALID_ARG;
  }
  if (tn_device_count() <= 0) {
    g_last_error = "no HIP device visible";
    return TN_DEVICE;
  }
  tn_engine* e = new tn_engine();
  tn_status st = e->init(stream, chi_max);
  if (st != TN_OK) {
    delete e;
    return st;
  }
  *out = e;
  return TN_OK;
}

void tn_engine_free(tn_engine* e) {
  if (!e) return;
  if (e->users > 0) {
    e->auto_free = true;  // the last handle frees it
    return;
  }
  delete e;
}

tn_status tn_mps_new_qubits_in(tn_engine* e, size_t n, tn_trunc trunc, size_t chi_cap,
                               tn_mps** out) {
  if (!e || !out) return TN_INVALID_ARG;
  *out = nullptr;
  if (n == 0) return TN_EMPTY_SYSTEM;  // mps.rs:375
  TN_TRY(check_trunc(trunc));
  return new_handle(e, n, trunc, chi_cap, out);
}

tn_status tn_mps_new_qubits(size_t n, tn_trunc trunc, size_t chi_cap, void* stream,
                            tn_mps** out) {
  if (!out) return TN_INVALID_ARG;
  *out = nullptr;
  if (n == 0) return TN_EMPTY_SYSTEM;  // mps.rs:375
  TN_TRY(check_trunc(trunc));
  if (tn_device_count() <= 0) {
    g_last_error = "no HIP device visible";
    return TN_DEVICE;
  }
  const size_t cap = default_cap(n, trunc, chi_cap);
  size_t maxb = 1;
  for (size_t b = 0; b + 1 < n; ++b) maxb = std::max(maxb, std::min(cap, exact_bond_max(n, b)));
  tn_engine* e = nullptr;
  TN_TRY(tn_engine_new(maxb, stream, &e));
  tn_status st = new_handle(e, n, trunc, chi_cap, out);
  if (st != TN_OK) {
    delete e;
    return st;
  }
  e->auto_free = true;
  return TN_OK;
}

tn_status tn_mps_from_vector(size_t n, const double* psi, tn_trunc trunc, size_t chi_cap,
                             void* stream, tn_mps** out) {
  if (!out) return TN_INVALID_ARG;
  *out = nullptr;
  if (n == 0) return TN_EMPTY_SYSTEM;  // mps.rs:527
  if (!psi || n > 30) return TN_INVALID_ARG;
  TN_TRY(check_trunc(trunc));
  const size_t len = (size_t)1 << n;
  // normalise (mps.rs:531-537); a zero vector has no normalised form
  double nrm2 = 0.0;
  for (size_t i = 0; i < 2 * len; ++i) nrm2 += psi[i] * psi[i];
  if (!(nrm2 > 0.0) || !std::isfinite(nrm2)) {
    g_last_error = "tn_mps_from_vector: the state vector has no finite non-zero norm";
    return TN_INVALID_ARG;
  }
  const double nrm = std::sqrt(nrm2);
  tn_mps* h = nullptr;
  TN_TRY(tn_mps_new_qubits(n, trunc, chi_cap, stream, &h));
  tn_engine* e = h->eng;
  // the conjugated normalised vector: q = psi as 2 x 2^(n-1), row-major
  std::vector<double2> q0(len);
  for (size_t i = 0; i < len; ++i)
    q0[i] = make_double2(psi[2 * i] / nrm, -psi[2 * i + 1] / nrm);
  if (n == 1) {  // mps.rs:538-544: Gamma[0] = the vector
    std::vector<double2> g(2);
    for (int i = 0; i < 2; ++i) g[i] = make_double2(q0[i].x, -q0[i].y);
    const hipError_t er = hipMemcpy(h->gam, g.data(), sizeof(double2) * 2, hipMemcpyHostToDevice);
    if (er != hipSuccess) {
      tn_mps_free(h);
      g_last_error = hipGetErrorString(er);
      return TN_DEVICE;
    }
    *out = h;
    return TN_OK;
  }
  // temporaries of the call: two factor buffers, the SVD workspace of one
  // task with C <= kMaxCols columns, and its descriptor
  constexpr int CM = kMaxCols;
  const size_t gp = std::max(((size_t)CM / kBW / 2 + 1) * kMaxSplit * (size_t)(4 * kBW * kBW),
                             ((size_t)CM / kBWWide / 2 + 1) * kMaxSplit * (size_t)(4 * kBWWide * kBWWide));
  double2 *bufA = nullptr, *bufB = nullptr, *W = nullptr, *gpart = nullptr;
  double *sraw = nullptr, *sval = nullptr, *fac = nullptr, *part = nullptr, *offs = nullptr;
  int *perm = nullptr, *info = nullptr, *act = nullptr;
  UpdDesc* dd = nullptr;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(e->stream);
    hfree(bufA); hfree(bufB); hfree(W); hfree(gpart); hfree(sraw); hfree(sval); hfree(fac);
    hfree(part); hfree(offs); hfree(perm); hfree(info); hfree(act); hfree(dd);
  };
  auto fail = [&](tn_status st) {
    cleanup();
    tn_mps_free(h);
    return st;
  };
#define FV_HIP(expr)                                                          \
  do {                                                                        \
    hipError_t e__ = (expr);                                                  \
    if (e__ != hipSuccess) {                                                  \
      g_last_error = std::string(#expr) + ": " + hipGetErrorString(e__);      \
      return fail(TN_DEVICE);                                                 \
    }                                                                         \
  } while (0)
  FV_HIP(hipMalloc(&bufA, sizeof(double2) * len));
  FV_HIP(hipMalloc(&bufB, sizeof(double2) * len));
  FV_HIP(hipMalloc(&W, sizeof(double2) * (size_t)CM * CM));
  FV_HIP(hipMalloc(&gpart, sizeof(double2) * gp));
  FV_HIP(hipMalloc(&sraw, sizeof(double) * CM));
  FV_HIP(hipMalloc(&sval, sizeof(double) * CM));
  FV_HIP(hipMalloc(&fac, sizeof(double) * CM));
  FV_HIP(hipMalloc(&part, sizeof(double) * 2 * kNPart));
  FV_HIP(hipMalloc(&offs, sizeof(double) * kMaxSweeps));
  FV_HIP(hipMalloc(&perm, sizeof(int) * CM));
  FV_HIP(hipMalloc(&info, sizeof(int) * kInfoLen));
  FV_HIP(hipMalloc(&act, sizeof(int) * 2));
  FV_HIP(hipMalloc(&dd, sizeof(UpdDesc)));
  FV_HIP(hipMemcpyAsync(bufA, q0.data(), sizeof(double2) * len, hipMemcpyHostToDevice, e->stream));
  double h_offs[kMaxSweeps];
  double2* cur = bufA;  // conj(q), row-major m x ncols
  double2* oth = bufB;
  size_t udim = 1, qlen = len;
  for (size_t k = 0; k + 1 < n; ++k) {
    // mps.rs:465-470: fuse the site's index with the previous Schmidt index
    const size_t m = 2 * udim, ncols = qlen / m;
    const int orient = m <= ncols ? 0 : 1;
    FvArgs a{};
    a.m = (int)m;
    a.ncols = (int)ncols;
    a.R = (int)std::max(m, ncols);
    a.C = (int)std::min(m, ncols);
    a.orient = orient;
    a.src = cur;
    a.X = orient == 0 ? cur : oth;
    a.W = W;
    a.offs = offs;
    a.info = info;
    launch_fv_init(a, e->stream);
    FV_HIP(hipGetLastError());
    Workspace w;
    w.X = a.X; w.W = W; w.sraw = sraw; w.sval = sval; w.fac = fac; w.perm = perm;
    w.part = part; w.gpart = gpart; w.gpart_per = gp; w.gpart_slots = 1; w.offs = offs;
    w.info = info; w.act = act; w.h_offs = h_offs; w.ldx = a.R; w.ldw = a.C;
    const tn_tune& tu = e->tune_for((size_t)a.C);
    int sweeps = 0;
    size_t launches = 0;
    const int trace[3] = {0, 0, 0};
    const tn_status js = jacobi_sweeps(e, w, 1, (size_t)a.R, (size_t)a.C, tu,
                                       e->tol.for_prec(TN_PREC_F64), trace,
                                       [&](int) { h->flags |= kFlagNoConv; }, &sweeps, &launches);
    if (js != TN_OK) return fail(js);
    // process_svd (mps.rs:236-293) on the SVD: values, rank, lambda[k], dims[k+1]
    UpdDesc d;
    std::memset(&d, 0, sizeof(d));
    d.X = a.X; d.W = W; d.sraw = sraw; d.sval = sval; d.fac = fac; d.perm = perm;
    d.part = part; d.offs = offs; d.info = info; d.ldx = a.R; d.ldw = a.C;
    d.lamM = h->lam + h->loff[k];
    d.dims = h->dims + k;
    d.tkind = h->trunc.kind;
    d.tchi = (int)std::min<size_t>(h->trunc.chi, 1u << 30);
    d.teps = h->trunc.eps;
    d.cap_rank = (int)h->bmax[k];
    d.hflags = h->dflags();
    d.eflags = e->d_eflags;
    FV_HIP(hipMemcpyAsync(dd, &d, sizeof(UpdDesc), hipMemcpyHostToDevice, e->stream));
    launch_finish(dd, 1, sweeps, e->stream);
    FV_HIP(hipGetLastError());
    int rank = 0;
    FV_HIP(hipMemcpyAsync(&rank, h->dims + k + 1, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    FV_HIP(hipStreamSynchronize(e->stream));
    // Gamma[k] = U[:, :r] / lambda[k-1] (mps.rs:478-492); q <- the right factor
    a.rank = rank;
    a.perm = perm;
    a.sraw = sraw;
    a.fac = fac;
    a.gk = h->gam + h->goff[k];
    a.lamPrev = k == 0 ? nullptr : h->lam + h->loff[k - 1];
    a.X2 = orient == 0 ? oth : cur;
    launch_fv_split(a, e->stream);
    FV_HIP(hipGetLastError());
    h->ub[k] = (size_t)rank;
    cur = a.X2;
    oth = cur == bufA ? bufB : bufA;
    udim = (size_t)rank;
    qlen = udim * ncols;
  }
  // mps.rs:499-505: the last Gamma from the final factor
  launch_fv_last(cur, h->lam + h->loff[n - 2], (int)udim, h->gam + h->goff[n - 1], e->stream);
  FV_HIP(hipGetLastError());
#undef FV_HIP
  // lambda[k-1] Gamma[k] = U has orthonormal columns for k < n-1
  for (size_t k = 0; k + 1 < n; ++k) h->fresh[k] = kFreshLeft;
  h->fresh[n - 1] = 0;
  TN_TRY(fetch_eflags(e));
  const hipError_t er = hipStreamSynchronize(e->stream);
  if (er != hipSuccess) {
    g_last_error = hipGetErrorString(er);
    return fail(TN_DEVICE);
  }
  const tn_status st = collect_flags(e, &h, 1);
  cleanup();
  *out = h;
  return st;
}

void tn_mps_free(tn_mps* h) { delete h; }

tn_status tn_mps_clone(const tn_mps* h, tn_mps** out) {
  if (!h || !out) return TN_INVALID_ARG;
  tn_mps* c = nullptr;
  if (h->eng->auto_free) {
    TN_TRY(tn_mps_new_qubits(h->n, h->trunc, h->cap,
                             h->eng->own_stream ? nullptr : (void*)h->eng->stream, &c));
  } else {
    TN_TRY(tn_mps_new_qubits_in(h->eng, h->n, h->trunc, h->cap, &c));
  }
  c->prec = h->prec;
  TN_HIP(hipStreamSynchronize(h->eng->stream));
  TN_HIP(hipMemcpy(c->gam, h->gam, sizeof(double2) * h->gam_elems, hipMemcpyDeviceToDevice));
  TN_HIP(hipMemcpy(c->lam, h->lam, sizeof(double) * h->lam_elems, hipMemcpyDeviceToDevice));
  TN_HIP(hipMemcpy(c->dims, h->dims, sizeof(int) * (h->n + 1), hipMemcpyDeviceToDevice));
  c->ub = h->ub;
  c->fresh = h->fresh;
  *out = c;
  return TN_OK;
}

size_t tn_mps_n(const tn_mps* h) { return h ? h->n : 0; }

tn_status tn_get_trunc(const tn_mps* h, tn_trunc* out) {
  if (!h || !out) return TN_INVALID_ARG;
  *out = h->trunc;
  return TN_OK;
}

tn_status tn_set_trunc(tn_mps* h, tn_trunc trunc) {
  if (!h) return TN_INVALID_ARG;
  TN_TRY(check_trunc(trunc));
  h->trunc = trunc;
  return TN_OK;
}

tn_status tn_engine_set_tuning(tn_engine* e, const tn_tune* rows, size_t count) {
  if (!e || !rows || count == 0 || count > 16) return TN_INVALID_ARG;
  for (size_t i = 0; i < count; ++i) {
    const tn_tune& r = rows[i];
    if (r.c_max < 1 || (i > 0 && r.c_max <= rows[i - 1].c_max)) return TN_INVALID_ARG;
    if (r.theta_tile != 32 && r.theta_tile != 64) return TN_INVALID_ARG;
    if (r.bw != kBW && r.bw != kBWWide) return TN_INVALID_ARG;
    if (r.split_max < 1 || r.split_max > kMaxSplit) return TN_INVALID_ARG;
    if (r.split_rows < 64) return TN_INVALID_ARG;
    if (r.ahead < 1 || r.ahead > 3) return TN_INVALID_ARG;
    if (r.qr != 0 && r.qr != 1) return TN_INVALID_ARG;
  }
  e->tune.assign(rows, rows + count);
  return TN_OK;
}

tn_status tn_engine_get_tuning(tn_engine* e, tn_tune* rows, size_t cap, size_t* count) {
  if (!e || !count || (cap > 0 && !rows)) return TN_INVALID_ARG;
  *count = e->tune.size();
  for (size_t i = 0; i < cap && i < e->tune.size(); ++i) rows[i] = e->tune[i];
  return TN_OK;
}

tn_status tn_set_tuning(tn_mps* h, const tn_tune* rows, size_t count) {
  if (!h) return TN_INVALID_ARG;
  return tn_engine_set_tuning(h->eng, rows, count);
}

tn_status tn_set_precision(tn_mps* h, int prec) {
  if (!h || (prec != TN_PREC_F64 && prec != TN_PREC_F32 && prec != TN_PREC_BF16X3))
    return TN_INVALID_ARG;
  h->prec = prec;
  return TN_OK;
}

tn_status tn_sync(tn_mps* h) {
  if (!h) return TN_INVALID_ARG;
  TN_HIP(hipStreamSynchronize(h->eng->stream));
  return TN_OK;
}

tn_status tn_apply_unis(tn_mps* h, size_t m, const size_t* ks, const int* dims,
                        const double* ops) {
  if (!h || (m && (!ks || !ops))) return TN_INVALID_ARG;
  std::vector<std::vector<std::vector<GateOp>>> groups(1);
  TN_TRY(gate_groups(h, m, ks, dims, ops, &groups[0]));
  TN_TRY(exec_gates(h->eng, {h}, groups));
  return sync_collect(h->eng, {h});
}

tn_status tn_apply_u2_noalloc(tn_mps* h, size_t k, const double* op) {
  if (!h || !op) return TN_INVALID_ARG;
  if (h->n == 1 || k >= h->n - 1) return TN_OK;  // mps.rs:1335
  tn_engine* e = h->eng;
  // theta is at most (2 chiL) x (2 chiR) with the host's bounds on bonds
  // k-1 and k+1
  if (2 * h->uL(k) > (size_t)kScratchCols || 2 * h->uR(k + 1) > (size_t)kScratchCols) {
    g_last_error = "tn_apply_u2_noalloc: bonds next to the update above 256 (theta above "
                   "512 x 512, the static scratch bound); use tn_apply_u2";
    return TN_INVALID_ARG;
  }
  if (!e->scratch_ptr) {
    g_last_error = "tn_apply_u2_noalloc: no scratch address on this device";
    return TN_DEVICE;
  }
  double h_offs[kMaxSweeps];
  unsigned char* b = e->scratch_ptr;
  Workspace w;
  w.X = (double2*)(b + kScrX);
  w.W = (double2*)(b + kScrW);
  w.sraw = (double*)(b + kScrSraw);
  w.sval = (double*)(b + kScrSval);
  w.fac = (double*)(b + kScrFac);
  w.perm = (int*)(b + kScrPerm);
  w.part = (double*)(b + kScrPart);
  w.offs = (double*)(b + kScrOffs);
  w.info = (int*)(b + kScrInfo);
  w.act = (int*)(b + kScrAct);
  w.gpart = (double2*)(b + kScrGpart);
  w.gpart_per = kScrGpartElems;
  w.qrE = (double2*)(b + kScrQrE);
  w.qrT = (double2*)(b + kScrQrT);
  w.qrPerm = (int*)(b + kScrQrPerm);
  w.qr_tslot = kScrQrTElems;
  w.gpart_slots = 1;
  if (e->split_fused) {
    w.Xb = (double2*)(b + kScrXb);
    w.xb_slots = 1;
  }
  w.h_offs = h_offs;
  w.ldx = w.ldw = kScratchCols;
  w.desc = (UpdDesc*)(b + kScrDesc);
  Task t;
  t.h = h;
  t.k = k;
  t.has_gate = true;
  std::memcpy(t.gate, op, sizeof(double) * 32);
  TN_TRY(run_chunk(e, &t, 1, &w));
  TN_TRY(fetch_eflags(e));
  TN_HIP(hipStreamSynchronize(e->stream));
  return collect_flags(e, &h, 1);
}

tn_status tn_apply_u2_multi(tn_mps* h, size_t k, size_t count, const double* ops) {
  if (!h || (count && !ops)) return TN_INVALID_ARG;
  if (h->n == 1 || k >= h->n - 1) return TN_OK;  // mps.rs:1400
  // acc <- op . acc from the identity, in the listed order (mps.rs:1404-1410)
  double acc[32] = {};
  for (int i = 0; i < 4; ++i) acc[2 * (5 * i)] = 1.0;
  for (size_t q = 0; q < count; ++q) {
    const double* op = ops + 32 * q;
    double nxt[32];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        double re = 0.0, im = 0.0;
        for (int l = 0; l < 4; ++l) {
          const double ar = op[2 * (4 * i + l)], ai = op[2 * (4 * i + l) + 1];
          const double br = acc[2 * (4 * l + j)], bi = acc[2 * (4 * l + j) + 1];
          re += ar * br - ai * bi;
          im += ar * bi + ai * br;
        }
        nxt[2 * (4 * i + j)] = re;
        nxt[2 * (4 * i + j) + 1] = im;
      }
    std::memcpy(acc, nxt, sizeof(acc));
  }
  return tn_apply_u2(h, k, acc);
}

tn_status tn_apply_u1(tn_mps* h, size_t k, const double* op) {
  if (!h || !op) return TN_INVALID_ARG;
  int d = 2;
  return tn_apply_unis(h, 1, &k, &d, op);
}

tn_status tn_apply_u2(tn_mps* h, size_t k, const double* op) {
  if (!h || !op) return TN_INVALID_ARG;
  int d = 4;
  return tn_apply_unis(h, 1, &k, &d, op);
}

tn_status tn_probs(tn_mps* h, size_t k, double* out) {
  if (!h || !out) return TN_INVALID_ARG;
  if (k >= h->n) return TN_NONE;  // mps.rs:1142
  TN_TRY(h->eng->ensure_results(1));
  TN_TRY(run_measures(h->eng, {{h, k, -2, 0.0, 0}}));
  return read_results(h->eng, 0, 1, nullptr, nullptr, out, nullptr);
}

tn_status tn_prob(tn_mps* h, size_t k, size_t p, double* out) {
  if (!h || !out) return TN_INVALID_ARG;
  if (k >= h->n || p >= 2) return TN_NONE;  // mps.rs:1072
  double both[2];
  TN_TRY(tn_probs(h, k, both));
  *out = both[p];
  return TN_OK;
}

tn_status tn_measure_prob(tn_mps* h, size_t k, double r, int* outcome, double* prob) {
  if (!h) return TN_INVALID_ARG;
  if (k >= h->n) return TN_NONE;  // mps.rs:1483
  TN_TRY(h->eng->ensure_results(1));
  TN_TRY(run_measures(h->eng, {{h, k, -1, r, 0}}));
  TN_TRY(refactor_sweep(h));
  int o = 0, f = 0;
  double pr = 0.0;
  TN_TRY(read_results(h->eng, 0, 1, &o, &pr, nullptr, &f));
  if (outcome) *outcome = o;
  if (prob) *prob = pr;
  return first_status((f & 2) ? TN_SAMPLE_RANGE : TN_OK, collect_flags(h->eng, {h}));
}

tn_status tn_measure_postsel_prob(tn_mps* h, size_t k, size_t p, double* prob) {
  if (!h) return TN_INVALID_ARG;
  if (k >= h->n || p >= 2) return TN_NONE;  // mps.rs:1504
  TN_TRY(h->eng->ensure_results(1));
  TN_TRY(run_measures(h->eng, {{h, k, (int)p, 0.0, 0}}));
  TN_TRY(refactor_sweep(h));
  int f = 0;
  double pr = 0.0;
  TN_TRY(read_results(h->eng, 0, 1, nullptr, &pr, nullptr, &f));
  if (prob) *prob = pr;
  return first_status((f & 1) ? TN_ZERO_PROB : TN_OK, collect_flags(h->eng, {h}));
}

tn_status tn_meas_layer(tn_mps* h, size_t m, const size_t* ks, const int* forced,
                        const double* r, size_t* out_count, size_t* out_ks,
                        int* out_outcomes, double* out_probs) {
  if (!h || (m && !ks)) return TN_INVALID_ARG;
  TN_TRY(check_meas_args(m, forced, r));
  TN_TRY(h->eng->ensure_results((int)std::max<size_t>(m, 1)));
  std::vector<std::vector<Op>> ops(1);
  std::vector<size_t> done;
  int slot = 0;
  meas_schedule(h, m, ks, forced, r, &slot, &ops[0], &done);
  TN_TRY(exec_ops(h->eng, {h}, ops));
  const int cnt = (int)done.size();
  std::vector<int> flags(cnt), outs(cnt);
  std::vector<double> probs(cnt);
  TN_TRY(read_results(h->eng, 0, cnt, outs.data(), probs.data(), nullptr, flags.data()));
  if (out_count) *out_count = done.size();
  for (int i = 0; i < cnt; ++i) {
    if (out_ks) out_ks[i] = done[i];
    if (out_outcomes) out_outcomes[i] = outs[i];
    if (out_probs) out_probs[i] = probs[i];
  }
  return first_status(check_flags(flags), collect_flags(h->eng, {h}));
}

tn_status tn_apply_bilayer(tn_mps* h, size_t nu, const size_t* uks, const int* udims,
                           const double* ops, size_t nm, const size_t* mks, const int* forced,
                           const double* r, size_t* out_count, size_t* out_ks,
                           int* out_outcomes, double* out_probs) {
  if (!h) return TN_INVALID_ARG;
  // a clamped rank or an SVD at the sweep limit still writes the state:
  // the measurement layer runs and the status is reported with its own
  const tn_status su = tn_apply_unis(h, nu, uks, udims, ops);
  if (su != TN_OK && su != TN_CAPACITY && su != TN_NO_CONVERGENCE) return su;
  const tn_status sm =
      tn_meas_layer(h, nm, mks, forced, r, out_count, out_ks, out_outcomes, out_probs);
  return sm != TN_OK ? sm : su;
}

tn_status tn_entropy_ry_schmidt(tn_mps* h, double a, size_t b, double* out) {
  if (!h || !out) return TN_INVALID_ARG;
  if (h->n < 2 || b >= h->n - 1) return TN_NONE;  // mps.rs:1252
  tn_engine* e = h->eng;
  TN_TRY(e->ensure_scalars(1));
  TN_TRY(run_entropies(e, {{h->lam + h->loff[b], h->dims + b + 1, a, 0}}));
  TN_HIP(hipMemcpyAsync(out, e->d_scalar, sizeof(double), hipMemcpyDeviceToHost, e->stream));
  TN_HIP(hipStreamSynchronize(e->stream));
  return TN_OK;
}

tn_status tn_entropies(tn_engine* e, size_t count, tn_mps* const* hs, const size_t* bonds,
                       const double* alphas, double* out) {
  if (!e || (count && (!hs || !bonds || !alphas || !out))) return TN_INVALID_ARG;
  if (count == 0) return TN_OK;
  if (count > (size_t)INT32_MAX) return TN_INVALID_ARG;
  std::vector<EntReq> es;
  for (size_t i = 0; i < count; ++i) {
    tn_mps* h = hs[i];
    if (!h || h->eng != e) return TN_INVALID_ARG;
    if (h->n >= 2 && bonds[i] < h->n - 1)
      es.push_back({h->lam + h->loff[bonds[i]], h->dims + bonds[i] + 1, alphas[i], (int)i});
  }
  TN_TRY(e->ensure_scalars((int)count));
  TN_TRY(run_entropies(e, es));
  std::vector<double> tmp(count);
  TN_HIP(hipMemcpyAsync(tmp.data(), e->d_scalar, sizeof(double) * count, hipMemcpyDeviceToHost,
                        e->stream));
  TN_HIP(hipStreamSynchronize(e->stream));
  for (size_t i = 0; i < count; ++i) {
    tn_mps* h = hs[i];
    out[i] = (h->n >= 2 && bonds[i] < h->n - 1) ? tmp[i] : NAN;
  }
  return TN_OK;
}

tn_status tn_contract(tn_mps* h, double* out, size_t cap) {
  if (!h || !out) return TN_INVALID_ARG;
  const size_t n = h->n;
  if (n > 30) return TN_INVALID_ARG;
  const size_t len = (size_t)1 << n;
  if (cap < 2 * len) return TN_INVALID_ARG;
  tn_engine* e = h->eng;
  double2 *a = nullptr, *b = nullptr;
  auto done = [&](tn_status st) {
    (void)hipStreamSynchronize(e->stream);
    hfree(a);
    hfree(b);
    return st;
  };
  if (hipMalloc(&a, sizeof(double2) * len) != hipSuccess ||
      hipMalloc(&b, sizeof(double2) * len) != hipSuccess)
    return done(TN_DEVICE);
  const double2 one = make_double2(1.0, 0.0);
  if (hipMemcpyAsync(a, &one, sizeof(double2), hipMemcpyHostToDevice, e->stream) != hipSuccess)
    return done(TN_DEVICE);
  for (size_t k = 0; k < n; ++k) {
    const size_t P = (size_t)1 << k;
    const size_t ubR = k + 1 < n ? h->bmax[k] : 1;
    launch_contract(a, P, h->dims + k, h->gam + h->goff[k],
                    k == 0 ? nullptr : h->lam + h->loff[k - 1], 2 * P * ubR, b, e->stream);
    std::swap(a, b);
  }
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(out, a, sizeof(double2) * len, hipMemcpyDeviceToHost, e->stream) !=
          hipSuccess ||
      hipStreamSynchronize(e->stream) != hipSuccess)
    return done(TN_DEVICE);
  return done(TN_OK);
}

tn_status tn_matrix_part(tn_mps* h, size_t a, size_t b, double* out, size_t cap) {
  if (!h || !out) return TN_INVALID_ARG;
  const size_t n = h->n;
  b = std::min(b, n);  // mps.rs:1746
  if (a >= b || b - a > kMatrixPartMaxSites) return TN_INVALID_ARG;
  const size_t D = (size_t)1 << (b - a);
  if (cap < 2 * D * D) return TN_INVALID_ARG;
  tn_engine* e = h->eng;
  std::vector<int> d(n + 1);
  TN_HIP(hipMemcpyAsync(d.data(), h->dims, sizeof(int) * (n + 1), hipMemcpyDeviceToHost,
                        e->stream));
  TN_HIP(hipStreamSynchronize(e->stream));
  // T after site k holds chiL 2^(k-a+1) rows of d[k+1] entries
  const size_t chiL = (size_t)d[a];
  size_t most = chiL * chiL;
  for (size_t k = a; k < b; ++k) most = std::max(most, chiL * ((size_t)2 << (k - a)) * d[k + 1]);
  if (most > kMatrixPartMaxElems) return TN_INVALID_ARG;
  std::vector<double2> eye(chiL * chiL, make_double2(0.0, 0.0));
  for (size_t i = 0; i < chiL; ++i) eye[i * chiL + i] = make_double2(1.0, 0.0);
  double2 *x = nullptr, *y = nullptr, *rho = nullptr;
  auto done = [&](tn_status st) {
    (void)hipStreamSynchronize(e->stream);
    hfree(x);
    hfree(y);
    hfree(rho);
    return st;
  };
  if (hipMalloc(&x, sizeof(double2) * most) != hipSuccess ||
      hipMalloc(&y, sizeof(double2) * most) != hipSuccess ||
      hipMalloc(&rho, sizeof(double2) * D * D) != hipSuccess)
    return done(TN_DEVICE);
  if (hipMemcpyAsync(x, eye.data(), sizeof(double2) * eye.size(), hipMemcpyHostToDevice,
                     e->stream) != hipSuccess)
    return done(TN_DEVICE);
  // rows (alpha, s_a .. s_k): lambda[a-1] enters as the first site's left
  // weight on the identity, the bonds inside the range as their sites' left
  // weights (mps.rs:1778-1791)
  size_t P = chiL;
  for (size_t k = a; k < b; ++k) {
    launch_contract(x, P, h->dims + k, h->gam + h->goff[k],
                    k == 0 ? nullptr : h->lam + h->loff[k - 1], 2 * P * (size_t)d[k + 1], y,
                    e->stream);
    std::swap(x, y);
    P *= 2;
  }
  // the tracers' lambda^2 (mps.rs:1801-1826): bond a-1 through both factors'
  // lambda[a-1], bond b-1 as the weight of the summed right index
  launch_rho_part(x, chiL, D, (size_t)d[b], b < n ? h->lam + h->loff[b - 1] : nullptr, rho,
                  e->stream);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(out, rho, sizeof(double2) * D * D, hipMemcpyDeviceToHost, e->stream) !=
          hipSuccess ||
      hipStreamSynchronize(e->stream) != hipSuccess)
    return done(TN_DEVICE);
  return done(TN_OK);
}

tn_status tn_entropy_vn(tn_mps* h, size_t b, double* out) {
  return tn_entropy_ry_schmidt(h, 1.0, b, out);
}

tn_status tn_bond_dims(tn_mps* h, size_t* out) {
  if (!h || (!out && h->n > 1)) return TN_INVALID_ARG;
  if (h->n < 2) return TN_OK;
  std::vector<int> d(h->n + 1);
  TN_HIP(hipMemcpyAsync(d.data(), h->dims, sizeof(int) * (h->n + 1), hipMemcpyDeviceToHost,
                        h->eng->stream));
  TN_HIP(hipStreamSynchronize(h->eng->stream));
  for (size_t b = 0; b + 1 < h->n; ++b) out[b] = (size_t)d[b + 1];
  return TN_OK;
}

tn_status tn_max_bond_dim(tn_mps* h, size_t* out) {
  if (!h || !out) return TN_INVALID_ARG;
  if (h->n < 2) return TN_NONE;
  std::vector<size_t> d(h->n - 1);
  TN_TRY(tn_bond_dims(h, d.data()));
  *out = *std::max_element(d.begin(), d.end());
  return TN_OK;
}

tn_status tn_schmidt(tn_mps* h, size_t b, double* buf, size_t cap, size_t* len) {
  if (!h || !len) return TN_INVALID_ARG;
  if (h->n < 2 || b >= h->n - 1) return TN_NONE;
  std::vector<size_t> d(h->n - 1);
  TN_TRY(tn_bond_dims(h, d.data()));
  *len = d[b];
  const size_t cnt = std::min(cap, d[b]);
  if (cnt && buf)
    TN_HIP(hipMemcpy(buf, h->lam + h->loff[b], sizeof(double) * cnt, hipMemcpyDeviceToHost));
  return TN_OK;
}

tn_status tn_get_gamma(tn_mps* h, size_t k, double* buf, size_t cap, size_t* shape) {
  if (!h || !shape) return TN_INVALID_ARG;
  if (k >= h->n) return TN_NONE;
  std::vector<int> d(h->n + 1);
  TN_HIP(hipStreamSynchronize(h->eng->stream));
  TN_HIP(hipMemcpy(d.data(), h->dims, sizeof(int) * (h->n + 1), hipMemcpyDeviceToHost));
  shape[0] = (size_t)d[k];
  shape[1] = 2;
  shape[2] = (size_t)d[k + 1];
  const size_t tot = shape[0] * 2 * shape[2];
  if (buf && cap >= 2 * tot)
    TN_HIP(hipMemcpy(buf, h->gam + h->goff[k], sizeof(double2) * tot, hipMemcpyDeviceToHost));
  else if (buf)
    return TN_INVALID_ARG;
  return TN_OK;
}

tn_status tn_run_circuit(tn_mps* h, size_t depth, const size_t* nu, const size_t* uks,
                         const double* ops, const size_t* nm, const size_t* mks,
                         const int* forced, const double* r, size_t ent_bond, double* ent_out,
                         size_t* out_counts, int* out_outcomes, double* out_probs) {
  if (!h) return TN_INVALID_ARG;
  tn_engine* e = h->eng;
  size_t total_m = 0;
  for (size_t t = 0; t < depth; ++t) total_m += nm[t];
  TN_TRY(check_meas_args(total_m, forced, r));
  TN_TRY(e->ensure_results((int)std::max<size_t>(total_m, 1)));
  TN_TRY(e->ensure_scalars((int)depth + 1));
  const bool want_ent = ent_out && h->n >= 2 && ent_bond < h->n - 1;
  const double* lamb = want_ent ? h->lam + h->loff[ent_bond] : nullptr;
  const int* dimb = want_ent ? h->dims + ent_bond + 1 : nullptr;
  if (want_ent) TN_TRY(run_entropies(e, {{lamb, dimb, 1.0, 0}}));
  size_t uo = 0, mo = 0;
  int slot = 0;
  for (size_t t = 0; t < depth; ++t) {
    std::vector<int> dsz(nu[t], 4);
    std::vector<std::vector<std::vector<GateOp>>> groups(1);
    TN_TRY(gate_groups(h, nu[t], uks + uo, dsz.data(), ops + 32 * uo, &groups[0]));
    TN_TRY(exec_gates(e, {h}, groups));
    uo += nu[t];
    std::vector<std::vector<Op>> mops(1);
    std::vector<size_t> done;
    meas_schedule(h, nm[t], mks + mo, forced ? forced + mo : nullptr, r ? r + mo : nullptr,
                  &slot, &mops[0], &done);
    TN_TRY(exec_ops(e, {h}, mops));
    mo += nm[t];
    if (out_counts) out_counts[t] = done.size();
    if (want_ent) TN_TRY(run_entropies(e, {{lamb, dimb, 1.0, (int)t + 1}}));
  }
  std::vector<int> flags(slot);
  TN_TRY(read_results(e, 0, slot, out_outcomes, out_probs, nullptr, flags.data()));
  if (want_ent)
    TN_HIP(hipMemcpy(ent_out, e->d_scalar, sizeof(double) * (depth + 1), hipMemcpyDeviceToHost));
  return first_status(check_flags(flags), collect_flags(e, {h}));
}

tn_status tn_batch_bilayer(tn_engine* e, size_t count, tn_mps* const* hs, const size_t* nu,
                           const size_t* uks, const double* ops, const size_t* nm,
                           const size_t* mks, const int* forced, const double* r,
                           const size_t* ent_bonds, double* ent_out, size_t* out_counts,
                           int* out_outcomes, double* out_probs) {
  if (!e || (count && (!hs || !nu || !nm))) return TN_INVALID_ARG;
  std::vector<tn_mps*> hv(hs, hs + count);
  for (tn_mps* h : hv)
    if (!h || h->eng != e) return TN_INVALID_ARG;
  size_t total_m = 0;
  for (size_t i = 0; i < count; ++i) total_m += nm[i];
  TN_TRY(check_meas_args(total_m, forced, r));
  TN_TRY(e->ensure_results((int)std::max<size_t>(total_m, 1)));
  TN_TRY(e->ensure_scalars((int)std::max<size_t>(count, 1)));
  // gates: group j of every handle in one launch batch
  std::vector<std::vector<std::vector<GateOp>>> groups(count);
  size_t uo = 0;
  for (size_t i = 0; i < count; ++i) {
    std::vector<int> dsz(nu[i], 4);
    TN_TRY(gate_groups(hv[i], nu[i], uks + uo, dsz.data(), ops + 32 * uo, &groups[i]));
    uo += nu[i];
  }
  TN_TRY(exec_gates(e, hv, groups));
  // measurement schedules in lock step
  std::vector<std::vector<Op>> mops(count);
  std::vector<size_t> counts(count);
  size_t mo = 0;
  int slot = 0;
  for (size_t i = 0; i < count; ++i) {
    std::vector<size_t> done;
    meas_schedule(hv[i], nm[i], mks + mo, forced ? forced + mo : nullptr, r ? r + mo : nullptr,
                  &slot, &mops[i], &done);
    counts[i] = done.size();
    mo += nm[i];
  }
  return finish_layers(e, hv, mops, counts, slot, ent_bonds, ent_out, out_counts, out_outcomes,
                       out_probs);
}

tn_status tn_rng_words(uint64_t seed, uint64_t c1, uint64_t c2, uint64_t c3, size_t nblocks,
                       uint64_t* out) {
  if (nblocks && !out) return TN_INVALID_ARG;
  for (size_t b = 0; b < nblocks; ++b) {
    const Philox4x64 w = philox4x64_10(b, c1, c2, c3, seed, 0);
    for (int j = 0; j < 4; ++j) out[4 * b + j] = w.v[j];
  }
  return TN_OK;
}

tn_status tn_gen_haar4(tn_engine* e, uint64_t seed, size_t count, const uint64_t* ids,
                       double* out) {
  if (!e || (count && (!ids || !out))) return TN_INVALID_ARG;
  if (count == 0) return TN_OK;
  if (count > (size_t)INT32_MAX) return TN_INVALID_ARG;
  TN_TRY(e->ensure_gen(count));
  TN_HIP(hipMemcpyAsync(e->gen_ids, ids, sizeof(uint64_t) * 3 * count, hipMemcpyHostToDevice,
                        e->stream));
  launch_haar4(seed, e->gen_ids, (int)count, e->gen_u, e->stream);
  TN_HIP(hipGetLastError());
  TN_HIP(hipMemcpyAsync(out, e->gen_u, sizeof(double2) * 16 * count, hipMemcpyDeviceToHost,
                        e->stream));
  TN_HIP(hipStreamSynchronize(e->stream));
  return TN_OK;
}

tn_status tn_batch_bilayer_gen(tn_engine* e, size_t count, tn_mps* const* hs, uint64_t seed,
                               const uint64_t* traj_ids, const uint64_t* layers,
                               const double* p_meas,
                               const size_t* ent_bonds, double* ent_out, size_t* out_counts,
                               size_t* out_sites, int* out_outcomes, double* out_probs) {
  if (!e || (count && (!hs || !traj_ids || !layers || !p_meas))) return TN_INVALID_ARG;
  std::vector<tn_mps*> hv(hs, hs + count);
  for (tn_mps* h : hv)
    if (!h || h->eng != e) return TN_INVALID_ARG;
  // gates on the brickwork sites of the layer (TileQ2, circuit.rs:523-546):
  // generated on the device, applied from there
  std::vector<uint64_t> ids;
  std::vector<std::pair<size_t, size_t>> at;  // (handle, site)
  for (size_t i = 0; i < count; ++i)
    for (size_t k = layers[i] & 1; k + 1 < hv[i]->n; k += 2) {
      ids.insert(ids.end(), {(uint64_t)k, layers[i], traj_ids[i]});
      at.push_back({i, k});
    }
  if (at.size() > (size_t)INT32_MAX) return TN_INVALID_ARG;
  if (!at.empty()) {
    TN_TRY(e->ensure_gen(at.size()));
    TN_HIP(hipMemcpyAsync(e->gen_ids, ids.data(), sizeof(uint64_t) * ids.size(),
                          hipMemcpyHostToDevice, e->stream));
    launch_haar4(seed, e->gen_ids, (int)at.size(), e->gen_u, e->stream);
    TN_HIP(hipGetLastError());
    std::vector<Task> tasks(at.size());
    for (size_t j = 0; j < at.size(); ++j) {
      tasks[j].h = hv[at[j].first];
      tasks[j].k = at[j].second;
      tasks[j].has_gate = true;
      tasks[j].gdev = e->gen_u + 16 * j;
    }
    TN_TRY(run_updates(e, tasks));
  }
  // measurements: qubit q measured when its mask uniform is below p
  // (uniform_meas, circuit.rs:580-584); one draw per measured qubit in
  // increasing site order, the order apply_meas_prob_layer consumes draws
  std::vector<std::vector<Op>> mops(count);
  std::vector<size_t> counts(count);
  int slot = 0;
  size_t total = 0;
  std::vector<std::vector<size_t>> sites(count);
  std::vector<std::vector<double>> draws(count);
  for (size_t i = 0; i < count; ++i) {
    const size_t n = hv[i]->n;
    for (size_t q = 0; q < n; ++q) {
      const Philox4x64 w = philox4x64_10(q / 4, kMaskStream, layers[i], traj_ids[i], seed, 0);
      if (u01(w.v[q % 4]) < p_meas[i]) sites[i].push_back(q);
    }
    for (size_t j = 0; j < sites[i].size(); ++j) {
      const Philox4x64 w = philox4x64_10(j / 4, kDrawStream, layers[i], traj_ids[i], seed, 0);
      draws[i].push_back(u01(w.v[j % 4]));
    }
    total += sites[i].size();
  }
  TN_TRY(e->ensure_results((int)std::max<size_t>(total, 1)));
  TN_TRY(e->ensure_scalars((int)std::max<size_t>(count, 1)));
  size_t so = 0;
  for (size_t i = 0; i < count; ++i) {
    std::vector<size_t> done;
    meas_schedule(hv[i], sites[i].size(), sites[i].data(), nullptr, draws[i].data(), &slot,
                  &mops[i], &done);
    counts[i] = done.size();
    if (out_sites)
      for (size_t q : done) out_sites[so++] = q;
  }
  return finish_layers(e, hv, mops, counts, slot, ent_bonds, ent_out, out_counts, out_outcomes,
                       out_probs);
}

tn_status tn_engine_stats(tn_engine* e, size_t* updates, size_t* sweeps) {
  if (!e) return TN_INVALID_ARG;
  if (updates) *updates = (size_t)e->updates;
  if (sweeps) *sweeps = (size_t)e->sweeps;
  return TN_OK;
}

tn_status tn_engine_set_timing(tn_engine* e, int on) {
  if (!e) return TN_INVALID_ARG;
  TN_HIP(hipStreamSynchronize(e->stream));
  TN_TRY(e->perf_drain(true));
  e->timing = on != 0;
  return TN_OK;
}

tn_status tn_engine_get_perf(tn_engine* e, tn_perf* out) {
  if (!e || !out) return TN_INVALID_ARG;
  TN_TRY(e->perf_drain(true));
  *out = e->perf;
  out->ms_host_launch = e->perf_host_launch_ms;
  return TN_OK;
}

tn_status tn_engine_reset_perf(tn_engine* e) {
  if (!e) return TN_INVALID_ARG;
  TN_TRY(e->perf_drain(true));
  e->perf = tn_perf{};
  e->perf_host_launch_ms = 0.0;
  return TN_OK;
}

tn_status tn_engine_sync(tn_engine* e) {
  if (!e) return TN_INVALID_ARG;
  TN_HIP(hipStreamSynchronize(e->stream));
  return TN_OK;
}

tn_status tn_take_flags(tn_mps* h, int* flags) {
  if (!h || !flags) return TN_INVALID_ARG;
  *flags = h->flags;
  h->flags = 0;
  return TN_OK;
}

tn_status tn_stats(tn_mps* h, size_t* updates, size_t* sweeps) {
  if (!h) return TN_INVALID_ARG;
  if (updates) *updates = (size_t)h->eng->updates;
  if (sweeps) *sweeps = (size_t)h->eng->sweeps;
  return TN_OK;
}

tn_status tn_set_timing(tn_mps* h, int on) {
  if (!h) return TN_INVALID_ARG;
  TN_HIP(hipStreamSynchronize(h->eng->stream));
  TN_TRY(h->eng->perf_drain(true));
  h->eng->timing = on != 0;
  return TN_OK;
}

tn_status tn_get_perf(tn_mps* h, tn_perf* out) {
  if (!h || !out) return TN_INVALID_ARG;
  TN_TRY(h->eng->perf_drain(true));
  *out = h->eng->perf;
  out->ms_host_launch = h->eng->perf_host_launch_ms;
  return TN_OK;
}

tn_status tn_reset_perf(tn_mps* h) {
  if (!h) return TN_INVALID_ARG;
  TN_TRY(h->eng->perf_drain(true));
  h->eng->perf = tn_perf{};
  h->eng->perf_host_launch_ms = 0.0;
  return TN_OK;
}

}  // extern "C"
