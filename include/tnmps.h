/*
 * tnmps.h -- C ABI of the MI355X Schmidt-form (Gamma-Lambda) MPS update path.
 *
 * Drop-in boundary for tensor-net's `MPS<Q, C64>` operator API
 * (tensor-net/lib/mps.rs).  Every entry point below names the reference
 * method it replaces.  All state (Gamma tensors, Schmidt values, bond
 * dimensions) lives in device memory; calls are enqueued on the handle's HIP
 * stream and block only where they return a value to the host.
 *
 * Conventions carried over from the reference:
 *  - complex numbers are interleaved {re, im} doubles (num_complex::Complex64
 *    is #[repr(C)] {re, im});
 *  - operators are row-major; for two-qubit operators qubit k is the most
 *    significant bit of the 4x4 index (gate.rs:750-761);
 *  - a site index out of range is a silent no-op for unitaries (mps.rs:1301,
 *    1335) and TN_NONE ("None") for queries (mps.rs:1072, 1142, 1252, 1483);
 *  - out-of-range sites are dropped from measurement layers, the rest are
 *    sorted by site and deduplicated keeping the first (mps.rs:2212-2219);
 *  - one caller per handle at a time (the reference's &mut self); distinct
 *    handles are independent.
 */
#ifndef TNMPS_H
#define TNMPS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tn_mps tn_mps;
/* A HIP stream plus the update workspace shared by the handles made in it.
 * Handles of one engine can be advanced in lock step (tn_batch_bilayer). */
typedef struct tn_engine tn_engine;

/* Status codes.  1..5 mirror MPSError (mps.rs:126-154). */
typedef enum {
    TN_OK = 0,
    TN_EMPTY_SYSTEM = 1,                 /* MPSError::EmptySystem */
    TN_INVALID_QUANTUM_NUMBER = 2,       /* MPSError::InvalidQuantumNumber */
    TN_UNPHYSICAL_INDEX = 3,             /* MPSError::UnphysicalIndex */
    TN_STATE_INCOMPATIBLE_SHAPE = 4,     /* MPSError::StateIncompatibleShape */
    TN_OPERATOR_INCOMPATIBLE_SHAPE = 5,  /* MPSError::OperatorIncompatibleShape */
    TN_NONE = 6,          /* the reference returns None (index out of range) */
    TN_ZERO_PROB = 7,     /* a post-selected outcome had probability 0: the
                             state holds NaN/inf exactly as the reference's
                             does (mps.rs:1505-1507); reported, not prevented */
    TN_SAMPLE_RANGE = 8,  /* r >= total probability: the reference panics on
                             unwrap (mps.rs:1444); the state is projected on
                             the last outcome with non-zero probability */
    TN_CAPACITY = 9,      /* the rank rule asked for more values than the
                             bond's allocation (chi_cap) holds: the update
                             kept that many and the state was written; the
                             reference has no bound (None / Cutoff) */
    TN_DEVICE = 10,       /* HIP runtime error */
    TN_INVALID_ARG = 11,  /* null pointer / bad argument at the C boundary */
    TN_NO_CONVERGENCE = 12 /* an SVD stopped at its sweep limit before its
                              tolerance (the result is still written) */
} tn_status;

/* Bits of a handle's sticky failure flags (tn_take_flags).  Calls that run
 * updates return TN_CAPACITY / TN_NO_CONVERGENCE when any update they
 * enqueued set one (after any measurement status, which comes first); the
 * handles that set them keep the bits until tn_take_flags. */
enum { TN_FLAG_NO_CONVERGENCE = 1, TN_FLAG_CAPACITY = 2 };

/* BondDim (mps.rs:157-166).  kind 0 = None (Cutoff at f64::EPSILON),
 * 1 = Const(chi), 2 = Cutoff(eps). */
typedef struct {
    int kind;
    size_t chi;
    double eps;
} tn_trunc;

enum { TN_TRUNC_NONE = 0, TN_TRUNC_CONST = 1, TN_TRUNC_CUTOFF = 2 };

/* Arithmetic of the two-site contraction theta = U (lambda Gamma lambda Gamma
 * lambda): FP64; FP32 with FP32 accumulation; or each FP32 operand split into
 * two BF16 halves, three BF16 products per real product, FP32 accumulation.
 * The SVD, truncation, lambda renormalisation and every stored tensor stay
 * FP64 in all three modes. */
enum { TN_PREC_F64 = 0, TN_PREC_F32 = 1, TN_PREC_BF16X3 = 2 };

/* One row of an engine's launch-shape table.  An update batch whose largest
 * SVD column count C (min(m, n) of theta, or the middle bond K of a factored
 * update) is at most c_max uses the first such row; the last row also takes
 * larger C.  No counterpart in mps.rs: these choose launch shapes only, and
 * every choice computes the same SVD to rounding. */
typedef struct {
    int c_max;       /* bucket bound on C, rows in increasing order */
    int theta_tile;  /* 32 or 64: theta tile edge on the matrix cores */
    int bw;          /* 8 or 16: Jacobi block width (columns) */
    int split_max;   /* 1..8: most workgroups per block pair (row-split steps) */
    int split_rows;  /* >= 64: the row split S doubles while the rows are at
                        least split_rows x S (slices of split_rows / 2 rows or more);
                        a one-launch row-split step (a lone SVD or up to 4, at most
                        256 rows) doubles it further, up to split_max, while its
                        slices keep 32 rows and its launch one workgroup per CU */
    int ahead;       /* 1..3: sweeps enqueued ahead while far from convergence */
    int qr;          /* 0 or 1: precondition the SVD with a column-sorted
                        Householder QR and run the Jacobi sweeps on R^H
                        (taken when every SVD of the batch has at most 512
                        rows; the SVD computed is the same to rounding) */
} tn_tune;

/* Human-readable name of a status code (static string). */
const char* tn_status_str(tn_status s);

/* Text of the last HIP error seen by the calling thread (static storage). */
const char* tn_last_error(void);

/* Number of visible HIP devices (0 when none); does not create a context. */
int tn_device_count(void);

/* --------------------------------------------------------------- lifecycle */

/* MPS::new_qubits (mps.rs:2033-2039 -> MPS::new, mps.rs:370-391): n qubits in
 * |0...0>.  chi_cap bounds the device allocation per bond; 0 means
 * trunc.chi for Const, else the largest exact bond dimension of n qubits,
 * min(2^{b+1}, 2^{n-b-1}), up to 512.  A None / Cutoff state that needs more
 * than its bound keeps that many values and the call returns TN_CAPACITY.
 * stream is a hipStream_t or NULL (the library creates one).  n == 0 ->
 * TN_EMPTY_SYSTEM. */
tn_status tn_mps_new_qubits(size_t n, tn_trunc trunc, size_t chi_cap,
                            void* stream, tn_mps** out);
/* MPS::from_vector (mps.rs:516-546 -> factorize, mps.rs:448-508) for n
 * qubits: psi holds 2^n interleaved complex amplitudes, qubit 0 the slowest
 * index; it is normalised, then factored site by site, each step an SVD on
 * the device with process_svd's rank rule (trunc, chi_cap as in
 * tn_mps_new_qubits).  n == 0 -> TN_EMPTY_SYSTEM; a zero or non-finite
 * vector, or n > 30 -> TN_INVALID_ARG; a rank above chi_cap -> TN_CAPACITY
 * (the state is made). */
tn_status tn_mps_from_vector(size_t n, const double* psi, tn_trunc trunc, size_t chi_cap,
                             void* stream, tn_mps** out);
/* Drop (frees device memory).  A handle made by tn_mps_new_qubits frees its
 * private engine with it. */
void tn_mps_free(tn_mps* h);

/* An engine for handles whose bonds stay at or below chi_max (1..512), on
 * `stream` (hipStream_t, or NULL: the library creates one).  No reference
 * counterpart: the reference runs one MPS per thread; the engine lets many
 * trajectories share launches.  tn_engine_free defers the release until the
 * last handle of the engine is freed. */
tn_status tn_engine_new(size_t chi_max, void* stream, tn_engine** out);
void tn_engine_free(tn_engine* e);
/* MPS::new_qubits (mps.rs:2033-2039) inside an engine. */
tn_status tn_mps_new_qubits_in(tn_engine* e, size_t n, tn_trunc trunc, size_t chi_cap,
                               tn_mps** out);
/* Contraction arithmetic of this handle (TN_PREC_F64 default).  With
 * TN_PREC_F32 (TN_PREC_BF16X3) the rank rule drops normalised singular values
 * below 2^-20 (2^-13), and the Jacobi SVD stops after a sweep whose largest
 * relative off-diagonal is at most 1e-4 (1e-3) instead of 1e-8: theta's own
 * round-off is larger than what further sweeps would remove. */
tn_status tn_set_precision(tn_mps* h, int prec);
/* Replace the engine's launch-shape table (1..16 rows, c_max increasing,
 * fields in their ranges; TN_INVALID_ARG otherwise, table unchanged).  The
 * default table comes from the TNMPS_* environment knobs.  Takes effect from
 * the next update batch; the engine's caller holds it (one caller at a time). */
tn_status tn_engine_set_tuning(tn_engine* e, const tn_tune* rows, size_t count);
/* Copy the engine's table into rows[0 .. cap); *count = its row count. */
tn_status tn_engine_get_tuning(tn_engine* e, tn_tune* rows, size_t cap, size_t* count);
/* The same for the engine of a handle. */
tn_status tn_set_tuning(tn_mps* h, const tn_tune* rows, size_t count);
/* MPS::clone (#[derive(Clone)], mps.rs:337): deep copy on the same stream. */
tn_status tn_mps_clone(const tn_mps* h, tn_mps** out);
size_t tn_mps_n(const tn_mps* h);
/* get_trunc / set_trunc (mps.rs:1216-1221). */
tn_status tn_get_trunc(const tn_mps* h, tn_trunc* out);
tn_status tn_set_trunc(tn_mps* h, tn_trunc trunc);
/* Wait for all work enqueued on the handle's stream. */
tn_status tn_sync(tn_mps* h);

/* ---------------------------------------------------------------- unitaries */

/* MPS::apply_unitary1 (mps.rs:1298-1314).  op: 2x2 row-major, 8 doubles. */
tn_status tn_apply_u1(tn_mps* h, size_t k, const double* op);
/* MPS::apply_unitary2 (mps.rs:1332-1387): contract -> gate -> SVD ->
 * truncate -> split -> renormalise.  op: 4x4 row-major, 32 doubles. */
tn_status tn_apply_u2(tn_mps* h, size_t k, const double* op);
/* tn_apply_u2 without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory during the call (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same update, kernels and result as tn_apply_u2 (mps.rs:1332-1387), with
 * the update's workspace (X, W, per-column arrays, row-split partial Gram
 * matrices, the QR preconditioner's buffers, the descriptor) in a static
 * device array of 18.3 MiB per GPU,
 * held until the process ends, instead of the engine's.  Limit: the bonds
 * next to the update (k-1 and k+1) at most 256, so theta is at most
 * 512 x 512; above that TN_INVALID_ARG (tn_last_error says why) and the state
 * is untouched.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (from several host threads or engines) race.  Blocks
 * until the update is done.  The handle itself and its engine must exist
 * (their memory was allocated when they were made). */
tn_status tn_apply_u2_noalloc(tn_mps* h, size_t k, const double* op);
/* MPS::apply_unitary2_multi (mps.rs:1393-1411): count 4x4 operators (32
 * doubles each, back to back) multiplied in order, op[count-1] ... op[0],
 * then applied as one apply_unitary2 at bond k (count == 0 applies the
 * identity, which still refactors the bond, as the reference does).  No-op
 * for n == 1 or k >= n - 1. */
tn_status tn_apply_u2_multi(tn_mps* h, size_t k, size_t count, const double* ops);
/* A sequence of apply_uni calls (mps.rs:2093-2108) as in the unitary loop of
 * circuit::apply_bilayer (circuit.rs:637).  dims[i] is 2 or 4 (operator
 * side); ops holds the operators back to back (8 or 32 doubles each).
 * Consecutive two-qubit gates on disjoint site pairs run in one batched
 * launch; the result equals applying them one by one in the given order. */
tn_status tn_apply_unis(tn_mps* h, size_t m, const size_t* ks,
                        const int* dims, const double* ops);

/* ------------------------------------------------------------- measurement */

/* MPS::prob (mps.rs:1071-1074). */
tn_status tn_prob(tn_mps* h, size_t k, size_t p, double* out);
/* MPS::probs (mps.rs:1140-1143): out[2]. */
tn_status tn_probs(tn_mps* h, size_t k, double* out);
/* MPS::measure_prob (mps.rs:1479-1489) with the caller's uniform draw r in
 * [0,1) standing in for rng.gen::<f64>(); projection + full refactor sweep. */
tn_status tn_measure_prob(tn_mps* h, size_t k, double r, int* outcome,
                          double* prob);
/* MPS::measure_postsel_prob (mps.rs:1501-1510). */
tn_status tn_measure_postsel_prob(tn_mps* h, size_t k, size_t p, double* prob);
/* MPS::apply_meas_prob_layer (mps.rs:2206-2252).  Entry i measures site
 * ks[i]: forced[i] = 0/1 is Meas::Proj, -1 is Meas::Rand using draw r[i]
 * (r may be NULL if no entry is Rand).  Results for the processed entries
 * (sorted by site, deduplicated, out-of-range dropped) go to out_ks,
 * out_outcomes, out_probs (capacity m each); *out_count is their number. */
tn_status tn_meas_layer(tn_mps* h, size_t m, const size_t* ks,
                        const int* forced, const double* r, size_t* out_count,
                        size_t* out_ks, int* out_outcomes, double* out_probs);
/* circuit::apply_bilayer (circuit.rs:624-659): tn_apply_unis then
 * tn_meas_layer. */
tn_status tn_apply_bilayer(tn_mps* h, size_t nu, const size_t* uks,
                           const int* udims, const double* ops, size_t nm,
                           const size_t* mks, const int* forced,
                           const double* r, size_t* out_count, size_t* out_ks,
                           int* out_outcomes, double* out_probs);

/* ------------------------------------------------------------- observables */

/* MPS::entropy_vn (mps.rs:1251-1261). */
tn_status tn_entropy_vn(tn_mps* h, size_t b, double* out);
/* MPS::entropy_ry_schmidt (mps.rs:1268-1282). */
tn_status tn_entropy_ry_schmidt(tn_mps* h, double a, size_t b, double* out);
/* Schmidt values of bond b (svals[b], mps.rs:352-353): up to cap values. */
tn_status tn_schmidt(tn_mps* h, size_t b, double* buf, size_t cap,
                     size_t* len);
/* bond_dim for every bond (mps.rs:1243-1245): out[n-1]. */
tn_status tn_bond_dims(tn_mps* h, size_t* out);
/* max_bond_dim (mps.rs:1227-1229); TN_NONE when n == 1. */
tn_status tn_max_bond_dim(tn_mps* h, size_t* out);
/* Copy Gamma[k] (data[k], axes [u_{k-1}, s_k, u_k], row-major, interleaved
 * complex) into buf (cap doubles); shape[3] receives its dimensions. */
tn_status tn_get_gamma(tn_mps* h, size_t k, double* buf, size_t cap,
                       size_t* shape);

/* ------------------------------------------------------ trajectory driver */

/* Run `depth` bilayers (circuit.rs:624-659) on one handle with no host
 * round trip between layers.  Layer t has nu[t] two-qubit gates at sites
 * uks[..] with 4x4 operators ops[..] (32 doubles each) and nm[t]
 * measurements (mks, forced, r as in tn_meas_layer), all packed layer after
 * layer.  ent_bond: the bond whose entropy_vn is recorded before layer 1 and
 * after every layer into ent_out[depth+1] (may be NULL).  Outcomes and
 * probabilities of processed measurements are written to out_outcomes /
 * out_probs (capacity sum(nm)) in processing order; out_counts[t] receives
 * the number processed in layer t. */
tn_status tn_run_circuit(tn_mps* h, size_t depth, const size_t* nu,
                         const size_t* uks, const double* ops,
                         const size_t* nm, const size_t* mks,
                         const int* forced, const double* r, size_t ent_bond,
                         double* ent_out, size_t* out_counts,
                         int* out_outcomes, double* out_probs);

/* One bilayer (circuit::apply_bilayer, circuit.rs:624-659) on each of
 * `count` handles of engine e, in lock step: the gates of all handles run in
 * shared launches, then the measurement schedules (mps.rs:2206-2252) advance
 * together, step j of every handle in one launch.  Per handle i, in order:
 * nu[i] two-qubit gates (uks, ops: 32 doubles each) and nm[i] measurements
 * (mks, forced, r as in tn_meas_layer), packed handle after handle.
 * ent_bonds[i] (optional) selects the bond whose entropy_vn goes to
 * ent_out[i] after the layer; out_counts[i] receives the number of processed
 * measurements of handle i, whose outcomes and probabilities go, handle after
 * handle, to out_outcomes / out_probs.  Each handle's result equals running
 * tn_apply_bilayer on it alone. */
tn_status tn_batch_bilayer(tn_engine* e, size_t count, tn_mps* const* hs, const size_t* nu,
                           const size_t* uks, const double* ops, const size_t* nm,
                           const size_t* mks, const int* forced, const double* r,
                           const size_t* ent_bonds, double* ent_out, size_t* out_counts,
                           int* out_outcomes, double* out_probs);

/* ---- device-side workload generation ------------------------------------
 * Replaces the host generators of the reference's workloads: gate::haar
 * (gate.rs:888-909), haar_layer and uniform_meas (circuit.rs:558-584), with
 * a counter-based generator, Philox4x64-10 (Salmon et al., SC 2011), keyed by
 * (seed, 0).  A trajectory is fixed by (seed, trajectory id); its streams use
 * counter (block, c1, layer, trajectory id) with c1 = site for the Haar gate
 * on (site, site+1), c1 = 2^32 for the measurement mask, c1 = 2^33 for the
 * Born-sampling draws.  A 64-bit word w gives the uniform (w >> 11) 2^-53.
 * The random streams differ from the reference's rand/statrs streams; the
 * distributions are the same. */

/* Philox4x64-10 words of counters (b, c1, c2, c3), b = 0..nblocks-1, key
 * (seed, 0): 4 words per block into out. Host code, no device needed. */
tn_status tn_rng_words(uint64_t seed, uint64_t c1, uint64_t c2, uint64_t c3,
                       size_t nblocks, uint64_t* out);

/* `count` Haar-random 4x4 unitaries generated on the device, unitary u from
 * counter words ids[3u..3u+2] (blocks 0..7: 32 uniforms; Box-Muller pairs
 * give Re, Im of Z's entries row-major; Q of Z = QR with R's diagonal real
 * positive), copied to out (host, 32 doubles each, row-major interleaved). */
tn_status tn_gen_haar4(tn_engine* e, uint64_t seed, size_t count,
                       const uint64_t* ids, double* out);

/* One generated bilayer on each of `count` handles of engine e, in lock step
 * (as tn_batch_bilayer): layer t = layers[i] of trajectory traj_ids[i]
 * applies Haar gates generated on the device on the brickwork sites
 * k = t % 2, t % 2 + 2, ... < n - 1 (TileQ2, circuit.rs:523-546), then measures
 * each qubit q whose mask uniform is below p_meas[i] (uniform_meas,
 * circuit.rs:580-584) with apply_meas_prob_layer (mps.rs:2206-2252), the
 * j-th measured qubit sampled with the j-th draw.  out_sites receives the
 * measured qubits, handle after handle (out_counts[i] of them), alongside
 * out_outcomes / out_probs; ent_bonds / ent_out as in tn_batch_bilayer. No
 * gate or mask crosses the host-device boundary. */
tn_status tn_batch_bilayer_gen(tn_engine* e, size_t count, tn_mps* const* hs,
                               uint64_t seed, const uint64_t* traj_ids,
                               const uint64_t* layers, const double* p_meas,
                               const size_t* ent_bonds,
                               double* ent_out, size_t* out_counts, size_t* out_sites,
                               int* out_outcomes, double* out_probs);

/* ---- dense cross-checks and batched read-outs --------------------------- */

/* entropy_ry_schmidt (mps.rs:1268-1282; alphas[i] == 1 is entropy_vn,
 * :1251-1261) of bond bonds[i] of handle hs[i], for `count` read-outs in one
 * launch; NaN where the reference returns None (bond out of range). */
tn_status tn_entropies(tn_engine* e, size_t count, tn_mps* const* hs,
                       const size_t* bonds, const double* alphas, double* out);

/* MPS::contract (mps.rs:794-798): the flat state vector, qubit 0 slowest,
 * contracted on the device and copied to out (2^(n+1) doubles, interleaved;
 * cap is out's length in doubles).  n <= 30. */
tn_status tn_contract(tn_mps* h, double* out, size_t cap);

/* MPS::into_matrix_part(a..b) (mps.rs:1926-1943 over the network of
 * into_network_part, mps.rs:1742-1829): the reduced density matrix of sites
 * a..b-1, the others traced out through the canonical form (Gamma of the
 * range, lambda on its inner bonds, lambda^2 on bonds a-1 and b-1), as
 * 2^(b-a) x 2^(b-a) complex row-major in out (2^(2(b-a)+1) doubles,
 * interleaved; cap is out's length in doubles).  Site a is the most
 * significant bit of the row (ket) and column (bra) indices.  b is clamped
 * to n as the reference clamps part.end; a >= b, more than 12 sites, or a
 * contracted block of the range above 2^28 complex entries: TN_INVALID_ARG
 * (the reference has no bound). */
tn_status tn_matrix_part(tn_mps* h, size_t a, size_t b, double* out, size_t cap);

/* The handle's sticky failure flags (TN_FLAG_* bits) collected since the
 * last call, then cleared.  No reference counterpart: tells a batch caller
 * which of its handles had a rank clamped (TN_CAPACITY) or an SVD stopped at
 * the sweep limit (TN_NO_CONVERGENCE). */
tn_status tn_take_flags(tn_mps* h, int* flags);

/* Counters of the device work done by this handle (for throughput
 * accounting): number of two-site updates (contract -> SVD -> split) and
 * the sum over them of Jacobi sweeps. */
tn_status tn_stats(tn_mps* h, size_t* updates, size_t* sweeps);

/* Device-time accounting of one handle, enabled by tn_set_timing.  Times
 * are HIP-event spans on the handle's stream: ms_jacobi sums, per Jacobi
 * sweep, the span from the first to the last step kernel of that sweep.
 * alg_flops counts, per two-site update with theta of m x n (M = max, N =
 * min) and contracted bond K: 8mnK (contraction) + 32mn (gate, if any) +
 * svd part 4(4M^2N + 8MN^2 + 9N^3) (complex R-SVD with U and V). */
typedef struct {
    double ms_theta;        /* contraction (+ gate) and W initialisation */
    double ms_jacobi;       /* Jacobi SVD sweeps */
    double ms_post;         /* singular values, rank, split, renormalise */
    double ms_measure;      /* Born probabilities and projection */
    double alg_flops;       /* algorithmic flops of all updates */
    double svd_flops;       /* the SVD part of alg_flops */
    size_t updates;         /* two-site updates (gates and refactors) */
    size_t sweeps;          /* Jacobi sweeps summed over update batches */
    size_t jacobi_launches; /* Jacobi step kernels launched */
    double ms_host_launch;  /* host time spent enqueueing Jacobi step kernels */
    size_t factored;        /* updates that decomposed one factor of theta only */
    /* The memory-bound kernels: event spans and algorithmic bytes (each
     * Gamma element read once and written once; theta of m x n, new rank r):
     * split reads the factors and writes Gamma[k], Gamma[k+1] (32 r (m + n)
     * B), renormalise reads and writes both (32 r (m + n) B), a measurement
     * reads Gamma[k] (32 chiL chiR B) and, when it projects, writes it. */
    double ms_split;
    double ms_renorm;
    double bytes_split;
    double bytes_renorm;
    double bytes_measure;
    size_t measure_sites;
    /* The contraction kernel alone (k_theta: theta, gate, lambda weights; or
     * the factored updates' elementwise fill of X): its event span, and the
     * algorithmic flops of the contracted updates (8 m n K, + 32 m n with a
     * gate). */
    double ms_contract;
    double contract_flops;
    /* The QR preconditioner: SVD tasks that ran it (counted with timing on
     * or off), and the event span of its kernels (sort, QR, and the final
     * X <- Q (W Sigma); part of neither ms_theta nor ms_jacobi). */
    size_t qr_tasks;
    double ms_qr;
} tn_perf;

/* Turn device-time accounting on (1) or off (0).  On adds one stream
 * synchronisation per update batch. */
tn_status tn_set_timing(tn_mps* h, int on);
tn_status tn_get_perf(tn_mps* h, tn_perf* out);
tn_status tn_reset_perf(tn_mps* h);

/* The same counters and accounting for a whole engine (all its handles). */
tn_status tn_engine_stats(tn_engine* e, size_t* updates, size_t* sweeps);
tn_status tn_engine_set_timing(tn_engine* e, int on);
tn_status tn_engine_get_perf(tn_engine* e, tn_perf* out);
tn_status tn_engine_reset_perf(tn_engine* e);
/* Wait for all work enqueued on the engine's stream. */
tn_status tn_engine_sync(tn_engine* e);

#ifdef __cplusplus
}
#endif
#endif /* TNMPS_H */
