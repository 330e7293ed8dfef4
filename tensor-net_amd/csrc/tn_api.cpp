// Host driver and C ABI of the MI355X Schmidt-form MPS update path.
//
// The driver restates the control flow of tensor-net's MPS methods
// (tensor-net/lib/mps.rs) -- which bonds are refactored in which order, when
// sites are measured, what is a no-op -- and enqueues the device work for
// each step on the handle's HIP stream.  All numerical work happens in the
// kernels of tn_kernels.hip; Gamma/lambda never leave the device except when
// the caller asks for them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tnmps.h"
#include "tn_device.h"
#include "tn_kernels.h"

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

double env_double(const char* name, double dflt) {
  const char* v = std::getenv(name);
  if (!v || !*v) return dflt;
  return std::atof(v);
}

// Jacobi tolerances.  A pair of columns is rotated when |g_ij| exceeds
// tol_rot * sqrt(g_ii g_jj); the SVD has converged after the first sweep whose
// largest such ratio is at most conv_tol (Jacobi converges quadratically, so
// the residual after that sweep is of order conv_tol^2).
struct Tolerances {
  double tol_rot = 1e-15;
  double conv_tol = 1e-10;
  Tolerances() {
    tol_rot = env_double("TNMPS_TOL_ROT", tol_rot);
    conv_tol = env_double("TNMPS_CONV_TOL", conv_tol);
  }
};

void hfree(void* p) {
  if (p) (void)hipFree(p);
}
void hhfree(void* p) {
  if (p) (void)hipHostFree(p);
}

const Tolerances& tolerances() {
  static Tolerances t;
  return t;
}

}  // namespace

// --------------------------------------------------------------------------
// Engine: workspace for concurrent update tasks plus staging buffers.
// --------------------------------------------------------------------------
struct Engine {
  int dev = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int ldx = 0, ldw = 0;  // row capacities of X (R) and W (C)
  int slots = 0;         // concurrent update tasks the workspace holds
  double2* X = nullptr;
  double2* W = nullptr;
  double* sraw = nullptr;
  double* sval = nullptr;
  double* fac = nullptr;
  int* perm = nullptr;
  double* part = nullptr;
  double* offs = nullptr;
  int* info = nullptr;
  UpdDesc* d_desc = nullptr;
  UpdDesc* h_desc = nullptr;
  double* h_offs = nullptr;
  hipEvent_t desc_ev = nullptr;
  // measurement staging and results
  int mslots = 0;
  MeasDesc* d_mdesc = nullptr;
  MeasDesc* h_mdesc = nullptr;
  hipEvent_t mdesc_ev = nullptr;
  int* d_out = nullptr;
  double* d_prob = nullptr;
  double* d_probs = nullptr;
  int* d_flag = nullptr;
  double* d_scalar = nullptr;  // scalar results (entropies)
  int nscalar = 0;
  // statistics
  unsigned long long updates = 0, sweeps = 0;
  bool timing = false;
  hipEvent_t tev[6] = {};
  tn_perf perf{};
  std::vector<int> h_info;

  ~Engine() { release(); }

  void release_ws() {
    hfree(X); hfree(W); hfree(sraw); hfree(sval); hfree(fac); hfree(perm);
    hfree(part); hfree(offs); hfree(info); hfree(d_desc);
    hhfree(h_desc);
    hhfree(h_offs);
    X = W = nullptr; sraw = sval = fac = part = offs = nullptr; perm = info = nullptr;
    d_desc = nullptr; h_desc = nullptr; h_offs = nullptr;
    slots = 0;
  }
  void release_meas() {
    hfree(d_mdesc); hfree(d_out); hfree(d_prob); hfree(d_probs); hfree(d_flag);
    hhfree(h_mdesc);
    d_mdesc = nullptr; h_mdesc = nullptr; d_out = nullptr; d_prob = d_probs = nullptr;
    d_flag = nullptr; mslots = 0;
  }
  void release() {
    if (stream) (void)hipStreamSynchronize(stream);
    release_ws();
    release_meas();
    hfree(d_scalar); d_scalar = nullptr; nscalar = 0;
    for (hipEvent_t& ev : tev)
      if (ev) (void)hipEventDestroy(ev), ev = nullptr;
    if (desc_ev) (void)hipEventDestroy(desc_ev);
    if (mdesc_ev) (void)hipEventDestroy(mdesc_ev);
    desc_ev = mdesc_ev = nullptr;
    if (own_stream && stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }

  tn_status init(int device, void* s, int ldx_, int ldw_) {
    dev = device;
    TN_HIP(hipSetDevice(dev));
    if (s) {
      stream = (hipStream_t)s;
      own_stream = false;
    } else {
      TN_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
      own_stream = true;
    }
    ldx = ldx_;
    ldw = ldw_;
    TN_HIP(hipEventCreateWithFlags(&desc_ev, hipEventDisableTiming));
    TN_HIP(hipEventCreateWithFlags(&mdesc_ev, hipEventDisableTiming));
    for (hipEvent_t& ev : tev) TN_HIP(hipEventCreate(&ev));
    return TN_OK;
  }

  float span(int a, int b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, tev[a], tev[b]) != hipSuccess) return 0.f;
    return ms;
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
    TN_HIP(hipMalloc(&info, sizeof(int) * kInfoLen * ns));
    TN_HIP(hipMalloc(&d_desc, sizeof(UpdDesc) * ns));
    TN_HIP(hipHostMalloc(&h_desc, sizeof(UpdDesc) * ns, hipHostMallocDefault));
    TN_HIP(hipHostMalloc(&h_offs, sizeof(double) * ns, hipHostMallocDefault));
    slots = ns;
    return TN_OK;
  }

  tn_status ensure_mslots(int need) {
    if (need <= mslots) return TN_OK;
    TN_HIP(hipStreamSynchronize(stream));
    release_meas();
    const int ns = std::max(need, 64);
    TN_HIP(hipMalloc(&d_mdesc, sizeof(MeasDesc) * ns));
    TN_HIP(hipHostMalloc(&h_mdesc, sizeof(MeasDesc) * ns, hipHostMallocDefault));
    TN_HIP(hipMalloc(&d_out, sizeof(int) * ns));
    TN_HIP(hipMalloc(&d_prob, sizeof(double) * ns));
    TN_HIP(hipMalloc(&d_probs, sizeof(double) * 2 * ns));
    TN_HIP(hipMalloc(&d_flag, sizeof(int) * ns));
    mslots = ns;
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
  Engine* eng = nullptr;
  size_t n = 0;
  tn_trunc trunc{};
  size_t cap = 0;
  std::vector<size_t> bmax;     // allocation bound per bond
  std::vector<size_t> goff;     // element offset of each site slot
  std::vector<size_t> loff;     // element offset of each bond slot
  double2* gam = nullptr;
  double* lam = nullptr;
  int* dims = nullptr;
  std::vector<size_t> ub;       // host-side upper bound of each bond dimension
  size_t gam_elems = 0, lam_elems = 0;

  ~tn_mps() {
    if (eng) (void)hipStreamSynchronize(eng->stream);
    hfree(gam);
    hfree(lam);
    hfree(dims);
    delete eng;
  }
  size_t bL(size_t k) const { return k == 0 ? 1 : bmax[k - 1]; }
  size_t bR(size_t k) const { return k + 1 >= n ? 1 : bmax[k]; }
  size_t uL(size_t k) const { return k == 0 ? 1 : ub[k - 1]; }
  size_t uR(size_t k) const { return k + 1 >= n ? 1 : ub[k]; }
};

namespace {

size_t exact_bond_max(size_t n, size_t b) {
  // min(2^{b+1}, 2^{n-b-1}) without overflow
  const size_t l = b + 1, r = n - b - 1;
  const size_t e = std::min(l, r);
  if (e >= 40) return (size_t)1 << 40;
  return (size_t)1 << e;
}

size_t rank_bound(const tn_mps* h) {
  if (h->trunc.kind == TN_TRUNC_CONST) return std::max<size_t>(h->trunc.chi, 1);
  return (size_t)-1;
}

struct UpdReq {
  size_t k;
  bool has_gate;
  double gate[32];
};

// Enqueue a batch of independent two-site updates on one handle: bonds must
// pairwise differ by at least 2 so that no update reads what another writes.
tn_status run_updates(tn_mps* h, const std::vector<UpdReq>& reqs) {
  Engine* e = h->eng;
  const int nt = (int)reqs.size();
  if (nt == 0) return TN_OK;
  tn_status st = e->ensure_slots(nt);
  if (st != TN_OK) return st;
  TN_HIP(hipEventSynchronize(e->desc_ev));
  size_t maxC = 1, maxTiles = 1, maxCC = 1;
  for (int t = 0; t < nt; ++t) {
    const size_t k = reqs[t].k;
    UpdDesc& d = e->h_desc[t];
    std::memset(&d, 0, sizeof(d));
    d.gk = h->gam + h->goff[k];
    d.gk1 = h->gam + h->goff[k + 1];
    d.lamL = k == 0 ? nullptr : h->lam + h->loff[k - 1];
    d.lamM = h->lam + h->loff[k];
    d.lamR = (k + 2 >= h->n) ? nullptr : h->lam + h->loff[k + 1];
    d.dims = h->dims + k;
    d.X = e->X + (size_t)t * e->ldx * e->ldw;
    d.W = e->W + (size_t)t * e->ldw * e->ldw;
    d.sraw = e->sraw + (size_t)t * e->ldw;
    d.sval = e->sval + (size_t)t * e->ldw;
    d.fac = e->fac + (size_t)t * e->ldw;
    d.perm = e->perm + (size_t)t * e->ldw;
    d.part = e->part + (size_t)t * 2 * kNPart;
    d.offs = e->offs + t;
    d.info = e->info + (size_t)t * kInfoLen;
    d.ldx = e->ldx;
    d.ldw = e->ldw;
    d.has_gate = reqs[t].has_gate ? 1 : 0;
    d.tkind = h->trunc.kind;
    d.tchi = (int)std::min<size_t>(h->trunc.chi, 1u << 30);
    d.cap_rank = (int)h->bmax[k];
    d.teps = h->trunc.eps;
    if (reqs[t].has_gate)
      for (int i = 0; i < 16; ++i) d.gate[i] = make_double2(reqs[t].gate[2 * i], reqs[t].gate[2 * i + 1]);
    const size_t chiL = h->uL(k), chiR = h->uR(k + 1);
    const size_t C = std::min(2 * chiL, 2 * chiR);
    maxC = std::max(maxC, C);
    maxCC = std::max(maxCC, C * C);
    maxTiles = std::max(maxTiles, ((chiL + 3) / 4) * ((chiR + 63) / 64));
  }
  TN_HIP(hipMemcpyAsync(e->d_desc, e->h_desc, sizeof(UpdDesc) * nt, hipMemcpyHostToDevice,
                        e->stream));
  TN_HIP(hipEventRecord(e->desc_ev, e->stream));
  if (e->timing) TN_HIP(hipEventRecord(e->tev[0], e->stream));
  const int gx_theta = (int)std::min<size_t>(maxTiles, 1024);
  const int gx_w = (int)std::min<size_t>((maxCC + 255) / 256, 1024);
  launch_theta(e->d_desc, nt, gx_theta, e->stream);
  launch_winit(e->d_desc, nt, gx_w, e->stream);
  TN_HIP(hipGetLastError());
  if (e->timing) TN_HIP(hipEventRecord(e->tev[1], e->stream));
  const int nb = (int)((maxC + kBW - 1) / kBW);
  const int nbe = nb + (nb & 1);
  const int npairs = nbe / 2, steps = nbe - 1;
  const Tolerances& tol = tolerances();
  int sweeps_run = 0;
  for (int sw = 0; sw < kMaxSweeps; ++sw) {
    if (e->timing) TN_HIP(hipEventRecord(e->tev[2], e->stream));
    for (int s = 0; s < steps; ++s)
      launch_jacobi_step(e->d_desc, nt, npairs, sw, s, tol.tol_rot, tol.conv_tol, e->stream);
    TN_HIP(hipGetLastError());
    if (e->timing) TN_HIP(hipEventRecord(e->tev[3], e->stream));
    sweeps_run = sw + 1;
    TN_HIP(hipMemcpyAsync(e->h_offs, e->offs + (size_t)sw * nt, sizeof(double) * nt,
                          hipMemcpyDeviceToHost, e->stream));
    TN_HIP(hipStreamSynchronize(e->stream));
    if (e->timing) {
      if (sw == 0) e->perf.ms_theta += e->span(0, 1);
      e->perf.ms_jacobi += e->span(2, 3);
      e->perf.jacobi_launches += steps;
    }
    bool done = true;
    for (int t = 0; t < nt; ++t)
      if (!(e->h_offs[t] <= tol.conv_tol)) done = false;
    if (done) break;
  }
  if (e->timing) TN_HIP(hipEventRecord(e->tev[4], e->stream));
  launch_finish(e->d_desc, nt, sweeps_run, e->stream);
  launch_split(e->d_desc, nt, e->stream);
  launch_renorm(e->d_desc, nt, e->stream);
  TN_HIP(hipGetLastError());
  e->updates += nt;
  e->sweeps += (unsigned long long)sweeps_run * nt;
  if (e->timing) {
    TN_HIP(hipEventRecord(e->tev[5], e->stream));
    e->h_info.resize((size_t)nt * kInfoLen);
    TN_HIP(hipMemcpyAsync(e->h_info.data(), e->info, sizeof(int) * kInfoLen * nt,
                          hipMemcpyDeviceToHost, e->stream));
    TN_HIP(hipStreamSynchronize(e->stream));
    e->perf.ms_post += e->span(4, 5);
    e->perf.updates += nt;
    e->perf.sweeps += sweeps_run;
    for (int t = 0; t < nt; ++t) {
      const int* inf = e->h_info.data() + (size_t)t * kInfoLen;
      const double m = inf[kInfoM], n = inf[kInfoN], K = inf[kInfoChiM];
      const double M = std::max(m, n), N = std::min(m, n);
      const double fsvd = 4.0 * (4.0 * M * M * N + 8.0 * M * N * N + 9.0 * N * N * N);
      e->perf.svd_flops += fsvd;
      e->perf.alg_flops += 8.0 * m * n * K + (reqs[t].has_gate ? 32.0 * m * n : 0.0) + fsvd;
    }
  }
  const size_t rb = rank_bound(h);
  for (int t = 0; t < nt; ++t) {
    const size_t k = reqs[t].k;
    const size_t chiL = h->uL(k), chiR = h->uR(k + 1);
    h->ub[k] = std::min({2 * chiL, 2 * chiR, rb, h->bmax[k]});
  }
  return TN_OK;
}

tn_status refactor(tn_mps* h, size_t k) {
  UpdReq r;
  r.k = k;
  r.has_gate = false;
  return run_updates(h, {r});
}

// One site measurement; results land in slot `slot` of the engine's arrays.
tn_status enqueue_measure(tn_mps* h, size_t k, int forced, double r, int slot) {
  Engine* e = h->eng;
  tn_status st = e->ensure_mslots(slot + 1);
  if (st != TN_OK) return st;
  TN_HIP(hipEventSynchronize(e->mdesc_ev));
  MeasDesc& d = e->h_mdesc[0];
  std::memset(&d, 0, sizeof(d));
  d.g = h->gam + h->goff[k];
  d.lamL = k == 0 ? nullptr : h->lam + h->loff[k - 1];
  d.lamR = (k + 1 >= h->n) ? nullptr : h->lam + h->loff[k];
  d.dims = h->dims + k;
  d.forced = forced;
  d.slot = slot;
  d.r = r;
  d.out_outcome = e->d_out;
  d.out_prob = e->d_prob;
  d.out_probs = e->d_probs;
  d.out_flag = e->d_flag;
  TN_HIP(hipMemcpyAsync(e->d_mdesc, e->h_mdesc, sizeof(MeasDesc), hipMemcpyHostToDevice,
                        e->stream));
  TN_HIP(hipEventRecord(e->mdesc_ev, e->stream));
  if (e->timing) TN_HIP(hipEventRecord(e->tev[0], e->stream));
  launch_measure(e->d_mdesc, 1, e->stream);
  TN_HIP(hipGetLastError());
  if (e->timing) {
    TN_HIP(hipEventRecord(e->tev[1], e->stream));
    TN_HIP(hipStreamSynchronize(e->stream));
    e->perf.ms_measure += e->span(0, 1);
  }
  return TN_OK;
}

tn_status refactor_sweep(tn_mps* h) {
  // mps.rs:894-897
  if (h->n < 2) return TN_OK;
  for (size_t k = 0; k + 1 < h->n; ++k) {
    tn_status st = refactor(h, k);
    if (st != TN_OK) return st;
  }
  for (size_t k = h->n - 2; k-- > 0;) {
    tn_status st = refactor(h, k);
    if (st != TN_OK) return st;
  }
  return TN_OK;
}

tn_status read_meas(tn_mps* h, int count, int* outcomes, double* probs, double* both,
                    int* flags) {
  Engine* e = h->eng;
  if (count <= 0) return TN_OK;
  if (outcomes)
    TN_HIP(hipMemcpyAsync(outcomes, e->d_out, sizeof(int) * count, hipMemcpyDeviceToHost,
                          e->stream));
  if (probs)
    TN_HIP(hipMemcpyAsync(probs, e->d_prob, sizeof(double) * count, hipMemcpyDeviceToHost,
                          e->stream));
  if (both)
    TN_HIP(hipMemcpyAsync(both, e->d_probs, sizeof(double) * 2 * count,
                          hipMemcpyDeviceToHost, e->stream));
  if (flags)
    TN_HIP(hipMemcpyAsync(flags, e->d_flag, sizeof(int) * count, hipMemcpyDeviceToHost,
                          e->stream));
  TN_HIP(hipStreamSynchronize(e->stream));
  return TN_OK;
}

// apply_meas_prob_layer (mps.rs:2206-2252) without reading results back.
// Results of the processed entries go to engine slots base .. base+count-1.
tn_status meas_layer_enqueue(tn_mps* h, size_t m, const size_t* ks, const int* forced,
                             const double* r, int base, std::vector<size_t>* out_ks) {
  struct M {
    size_t k;
    int f;
    double r;
    size_t ord;
  };
  std::vector<M> ms;
  for (size_t i = 0; i < m; ++i) {
    if (ks[i] >= h->n) continue;  // mps.rs:2212-2216
    ms.push_back({ks[i], forced ? forced[i] : -1, r ? r[i] : 0.0, i});
  }
  out_ks->clear();
  if (ms.empty()) return TN_OK;  // mps.rs:2217
  std::stable_sort(ms.begin(), ms.end(), [](const M& a, const M& b) { return a.k < b.k; });
  std::vector<M> dd;
  for (const M& x : ms)
    if (dd.empty() || dd.back().k != x.k) dd.push_back(x);  // dedup_by_key keeps the first
  bool have_last = false;
  size_t last = 0;
  int slot = base;
  for (const M& x : dd) {
    if (have_last)
      for (size_t j = last; j < x.k; ++j) {  // mps.rs:2225-2228
        tn_status st = refactor(h, j);
        if (st != TN_OK) return st;
      }
    tn_status st = enqueue_measure(h, x.k, x.f >= 0 ? x.f : -1, x.r, slot++);
    if (st != TN_OK) return st;
    out_ks->push_back(x.k);
    have_last = true;
    last = x.k;
  }
  if (have_last)
    for (size_t j = last; j + 1 < h->n; ++j) {  // mps.rs:2246-2249
      tn_status st = refactor(h, j);
      if (st != TN_OK) return st;
    }
  if (h->n >= 2)
    for (size_t j = h->n - 2; j-- > 0;) {  // mps.rs:2250
      tn_status st = refactor(h, j);
      if (st != TN_OK) return st;
    }
  return TN_OK;
}

// Unitaries in order; consecutive two-qubit gates on bonds at least 2 apart
// run as one batch (they touch disjoint sites and bonds).
tn_status apply_unis_impl(tn_mps* h, size_t m, const size_t* ks, const int* dimsz,
                          const double* ops) {
  std::vector<UpdReq> batch;
  auto flush = [&]() -> tn_status {
    tn_status st = run_updates(h, batch);
    batch.clear();
    return st;
  };
  size_t off = 0;
  for (size_t i = 0; i < m; ++i) {
    const int dsz = dimsz ? dimsz[i] : 4;
    const double* op = ops + off;
    if (dsz == 2) {
      off += 8;
      if (ks[i] >= h->n) continue;  // mps.rs:1301
      tn_status st = flush();
      if (st != TN_OK) return st;
      launch_u1(h->gam + h->goff[ks[i]], h->dims + ks[i], op, 64, h->eng->stream);
      TN_HIP(hipGetLastError());
    } else if (dsz == 4) {
      off += 32;
      if (h->n == 1 || ks[i] >= h->n - 1) continue;  // mps.rs:1335
      bool conflict = false;
      for (const UpdReq& b : batch)
        if ((b.k > ks[i] ? b.k - ks[i] : ks[i] - b.k) < 2) conflict = true;
      if (conflict) {
        tn_status st = flush();
        if (st != TN_OK) return st;
      }
      UpdReq r;
      r.k = ks[i];
      r.has_gate = true;
      std::memcpy(r.gate, op, sizeof(double) * 32);
      batch.push_back(r);
    } else {
      tn_status st = flush();
      if (st != TN_OK) return st;
      return TN_OPERATOR_INCOMPATIBLE_SHAPE;  // mps.rs:2104
    }
  }
  return flush();
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

tn_status tn_mps_new_qubits(size_t n, tn_trunc trunc, size_t chi_cap, void* stream,
                            tn_mps** out) {
  if (!out) return TN_INVALID_ARG;
  *out = nullptr;
  if (n == 0) return TN_EMPTY_SYSTEM;  // mps.rs:375
  if (trunc.kind < 0 || trunc.kind > 2) return TN_INVALID_ARG;
  if (tn_device_count() <= 0) {
    g_last_error = "no HIP device visible";
    return TN_DEVICE;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  size_t cap = chi_cap;
  if (cap == 0) cap = trunc.kind == TN_TRUNC_CONST ? std::max<size_t>(trunc.chi, 1) : 256;
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
  if (2 * maxb > (size_t)kMaxCols) {
    // the singular-value sort of k_finish holds at most kMaxCols values
    g_last_error = "bond allocation above 512 (theta wider than 1024 columns)";
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
  h->eng = new Engine();
  tn_status st = h->eng->init(dev, stream, (int)(2 * maxb), (int)(2 * maxb));
  if (st != TN_OK) {
    delete h;
    return st;
  }
  auto fail = [&](hipError_t e) {
    g_last_error = hipGetErrorString(e);
    delete h;
    return TN_DEVICE;
  };
  hipError_t e;
  if ((e = hipMalloc(&h->gam, sizeof(double2) * h->gam_elems)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&h->lam, sizeof(double) * h->lam_elems)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&h->dims, sizeof(int) * (n + 1))) != hipSuccess) return fail(e);
  // |0...0>: Gamma[k] = (1,2,1) with [0,0,0] = 1, lambda = [1] (mps.rs:378-388)
  std::vector<double2> g0(h->gam_elems, make_double2(0.0, 0.0));
  for (size_t k = 0; k < n; ++k) g0[h->goff[k]] = make_double2(1.0, 0.0);
  std::vector<double> l0(h->lam_elems, 0.0);
  for (size_t b = 0; b + 1 < n; ++b) l0[h->loff[b]] = 1.0;
  std::vector<int> d0(n + 1, 1);
  Engine* en = h->eng;
  if ((e = hipMemcpyAsync(h->gam, g0.data(), sizeof(double2) * h->gam_elems,
                          hipMemcpyHostToDevice, en->stream)) != hipSuccess)
    return fail(e);
  if ((e = hipMemcpyAsync(h->lam, l0.data(), sizeof(double) * h->lam_elems,
                          hipMemcpyHostToDevice, en->stream)) != hipSuccess)
    return fail(e);
  if ((e = hipMemcpyAsync(h->dims, d0.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice,
                          en->stream)) != hipSuccess)
    return fail(e);
  if ((e = hipStreamSynchronize(en->stream)) != hipSuccess) return fail(e);
  *out = h;
  return TN_OK;
}

void tn_mps_free(tn_mps* h) { delete h; }

tn_status tn_mps_clone(const tn_mps* h, tn_mps** out) {
  if (!h || !out) return TN_INVALID_ARG;
  tn_status st = tn_mps_new_qubits(h->n, h->trunc, h->cap,
                                   h->eng->own_stream ? nullptr : (void*)h->eng->stream, out);
  if (st != TN_OK) return st;
  tn_mps* c = *out;
  TN_HIP(hipStreamSynchronize(h->eng->stream));
  TN_HIP(hipMemcpy(c->gam, h->gam, sizeof(double2) * h->gam_elems, hipMemcpyDeviceToDevice));
  TN_HIP(hipMemcpy(c->lam, h->lam, sizeof(double) * h->lam_elems, hipMemcpyDeviceToDevice));
  TN_HIP(hipMemcpy(c->dims, h->dims, sizeof(int) * (h->n + 1), hipMemcpyDeviceToDevice));
  c->ub = h->ub;
  return TN_OK;
}

size_t tn_mps_n(const tn_mps* h) { return h ? h->n : 0; }

tn_status tn_get_trunc(const tn_mps* h, tn_trunc* out) {
  if (!h || !out) return TN_INVALID_ARG;
  *out = h->trunc;
  return TN_OK;
}

tn_status tn_set_trunc(tn_mps* h, tn_trunc trunc) {
  if (!h || trunc.kind < 0 || trunc.kind > 2) return TN_INVALID_ARG;
  h->trunc = trunc;
  return TN_OK;
}

tn_status tn_sync(tn_mps* h) {
  if (!h) return TN_INVALID_ARG;
  TN_HIP(hipStreamSynchronize(h->eng->stream));
  return TN_OK;
}

tn_status tn_apply_u1(tn_mps* h, size_t k, const double* op) {
  if (!h || !op) return TN_INVALID_ARG;
  int d = 2;
  return apply_unis_impl(h, 1, &k, &d, op);
}

tn_status tn_apply_u2(tn_mps* h, size_t k, const double* op) {
  if (!h || !op) return TN_INVALID_ARG;
  int d = 4;
  return apply_unis_impl(h, 1, &k, &d, op);
}

tn_status tn_apply_unis(tn_mps* h, size_t m, const size_t* ks, const int* dims,
                        const double* ops) {
  if (!h || (m && (!ks || !ops))) return TN_INVALID_ARG;
  return apply_unis_impl(h, m, ks, dims, ops);
}

tn_status tn_probs(tn_mps* h, size_t k, double* out) {
  if (!h || !out) return TN_INVALID_ARG;
  if (k >= h->n) return TN_NONE;  // mps.rs:1142
  tn_status st = enqueue_measure(h, k, -2, 0.0, 0);
  if (st != TN_OK) return st;
  return read_meas(h, 1, nullptr, nullptr, out, nullptr);
}

tn_status tn_prob(tn_mps* h, size_t k, size_t p, double* out) {
  if (!h || !out) return TN_INVALID_ARG;
  if (k >= h->n || p >= 2) return TN_NONE;  // mps.rs:1072
  double both[2];
  tn_status st = tn_probs(h, k, both);
  if (st != TN_OK) return st;
  *out = both[p];
  return TN_OK;
}

tn_status tn_measure_prob(tn_mps* h, size_t k, double r, int* outcome, double* prob) {
  if (!h) return TN_INVALID_ARG;
  if (k >= h->n) return TN_NONE;  // mps.rs:1483
  tn_status st = enqueue_measure(h, k, -1, r, 0);
  if (st != TN_OK) return st;
  st = refactor_sweep(h);
  if (st != TN_OK) return st;
  int o = 0, f = 0;
  double pr = 0.0;
  st = read_meas(h, 1, &o, &pr, nullptr, &f);
  if (st != TN_OK) return st;
  if (outcome) *outcome = o;
  if (prob) *prob = pr;
  if (f & 2) return TN_SAMPLE_RANGE;
  return TN_OK;
}

tn_status tn_measure_postsel_prob(tn_mps* h, size_t k, size_t p, double* prob) {
  if (!h) return TN_INVALID_ARG;
  if (k >= h->n || p >= 2) return TN_NONE;  // mps.rs:1504
  tn_status st = enqueue_measure(h, k, (int)p, 0.0, 0);
  if (st != TN_OK) return st;
  st = refactor_sweep(h);
  if (st != TN_OK) return st;
  int f = 0;
  double pr = 0.0;
  st = read_meas(h, 1, nullptr, &pr, nullptr, &f);
  if (st != TN_OK) return st;
  if (prob) *prob = pr;
  if (f & 1) return TN_ZERO_PROB;
  return TN_OK;
}

tn_status tn_meas_layer(tn_mps* h, size_t m, const size_t* ks, const int* forced,
                        const double* r, size_t* out_count, size_t* out_ks,
                        int* out_outcomes, double* out_probs) {
  if (!h || (m && !ks)) return TN_INVALID_ARG;
  for (size_t i = 0; i < m; ++i) {
    if (forced && (forced[i] < -1 || forced[i] > 1)) return TN_INVALID_ARG;
    if ((!forced || forced[i] < 0) && !r) return TN_INVALID_ARG;
  }
  tn_status st = h->eng->ensure_mslots((int)std::max<size_t>(m, 1));
  if (st != TN_OK) return st;
  std::vector<size_t> done;
  st = meas_layer_enqueue(h, m, ks, forced, r, 0, &done);
  if (st != TN_OK) return st;
  const int cnt = (int)done.size();
  std::vector<int> flags(cnt);
  std::vector<int> outs(cnt);
  std::vector<double> probs(cnt);
  st = read_meas(h, cnt, outs.data(), probs.data(), nullptr, flags.data());
  if (st != TN_OK) return st;
  if (out_count) *out_count = done.size();
  for (int i = 0; i < cnt; ++i) {
    if (out_ks) out_ks[i] = done[i];
    if (out_outcomes) out_outcomes[i] = outs[i];
    if (out_probs) out_probs[i] = probs[i];
  }
  for (int i = 0; i < cnt; ++i)
    if (flags[i] & 2) return TN_SAMPLE_RANGE;
  for (int i = 0; i < cnt; ++i)
    if (flags[i] & 1) return TN_ZERO_PROB;
  return TN_OK;
}

tn_status tn_apply_bilayer(tn_mps* h, size_t nu, const size_t* uks, const int* udims,
                           const double* ops, size_t nm, const size_t* mks, const int* forced,
                           const double* r, size_t* out_count, size_t* out_ks,
                           int* out_outcomes, double* out_probs) {
  if (!h) return TN_INVALID_ARG;
  tn_status st = tn_apply_unis(h, nu, uks, udims, ops);
  if (st != TN_OK) return st;
  return tn_meas_layer(h, nm, mks, forced, r, out_count, out_ks, out_outcomes, out_probs);
}

tn_status tn_entropy_ry_schmidt(tn_mps* h, double a, size_t b, double* out) {
  if (!h || !out) return TN_INVALID_ARG;
  if (h->n < 2 || b >= h->n - 1) return TN_NONE;  // mps.rs:1252
  Engine* e = h->eng;
  tn_status st = e->ensure_scalars(1);
  if (st != TN_OK) return st;
  launch_entropy(h->lam + h->loff[b], h->dims + b + 1, a, e->d_scalar, e->stream);
  TN_HIP(hipGetLastError());
  TN_HIP(hipMemcpyAsync(out, e->d_scalar, sizeof(double), hipMemcpyDeviceToHost, e->stream));
  TN_HIP(hipStreamSynchronize(e->stream));
  return TN_OK;
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
  tn_status st = tn_bond_dims(h, d.data());
  if (st != TN_OK) return st;
  *out = *std::max_element(d.begin(), d.end());
  return TN_OK;
}

tn_status tn_schmidt(tn_mps* h, size_t b, double* buf, size_t cap, size_t* len) {
  if (!h || !len) return TN_INVALID_ARG;
  if (h->n < 2 || b >= h->n - 1) return TN_NONE;
  std::vector<size_t> d(h->n - 1);
  tn_status st = tn_bond_dims(h, d.data());
  if (st != TN_OK) return st;
  *len = d[b];
  const size_t cnt = std::min(cap, d[b]);
  if (cnt && buf) {
    TN_HIP(hipMemcpy(buf, h->lam + h->loff[b], sizeof(double) * cnt, hipMemcpyDeviceToHost));
  }
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
  Engine* e = h->eng;
  size_t total_m = 0;
  for (size_t t = 0; t < depth; ++t) total_m += nm[t];
  tn_status st = e->ensure_mslots((int)std::max<size_t>(total_m, 1));
  if (st != TN_OK) return st;
  st = e->ensure_scalars((int)depth + 1);
  if (st != TN_OK) return st;
  const bool want_ent = ent_out && h->n >= 2 && ent_bond < h->n - 1;
  if (want_ent)
    launch_entropy(h->lam + h->loff[ent_bond], h->dims + ent_bond + 1, 1.0, e->d_scalar,
                   e->stream);
  size_t uo = 0, mo = 0;
  int slot = 0;
  std::vector<size_t> done;
  for (size_t t = 0; t < depth; ++t) {
    std::vector<int> dsz(nu[t], 4);
    st = apply_unis_impl(h, nu[t], uks + uo, dsz.data(), ops + 32 * uo);
    if (st != TN_OK) return st;
    uo += nu[t];
    st = meas_layer_enqueue(h, nm[t], mks + mo, forced ? forced + mo : nullptr,
                            r ? r + mo : nullptr, slot, &done);
    if (st != TN_OK) return st;
    mo += nm[t];
    if (out_counts) out_counts[t] = done.size();
    slot += (int)done.size();
    if (want_ent)
      launch_entropy(h->lam + h->loff[ent_bond], h->dims + ent_bond + 1, 1.0,
                     e->d_scalar + t + 1, e->stream);
  }
  TN_HIP(hipGetLastError());
  std::vector<int> flags(slot);
  st = read_meas(h, slot, out_outcomes, out_probs, nullptr, flags.data());
  if (st != TN_OK) return st;
  if (want_ent) {
    TN_HIP(hipMemcpy(ent_out, e->d_scalar, sizeof(double) * (depth + 1), hipMemcpyDeviceToHost));
  }
  for (int i = 0; i < slot; ++i)
    if (flags[i] & 2) return TN_SAMPLE_RANGE;
  for (int i = 0; i < slot; ++i)
    if (flags[i] & 1) return TN_ZERO_PROB;
  return TN_OK;
}

tn_status tn_set_timing(tn_mps* h, int on) {
  if (!h) return TN_INVALID_ARG;
  TN_HIP(hipStreamSynchronize(h->eng->stream));
  h->eng->timing = on != 0;
  return TN_OK;
}

tn_status tn_get_perf(tn_mps* h, tn_perf* out) {
  if (!h || !out) return TN_INVALID_ARG;
  *out = h->eng->perf;
  return TN_OK;
}

tn_status tn_reset_perf(tn_mps* h) {
  if (!h) return TN_INVALID_ARG;
  h->eng->perf = tn_perf{};
  return TN_OK;
}

tn_status tn_stats(tn_mps* h, size_t* updates, size_t* sweeps) {
  if (!h) return TN_INVALID_ARG;
  if (updates) *updates = (size_t)h->eng->updates;
  if (sweeps) *sweeps = (size_t)h->eng->sweeps;
  return TN_OK;
}

}  // extern "C"
