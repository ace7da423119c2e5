/*
 * qrochet_hip.h -- C ABI of libqrochet_hip.so, the MI355X (gfx950) implementation
 * of Qrochet.jl's MPS (`Chain` ansatz) update path.
 *
 * Every entry point replaces one method of the reference's Julia API
 * (Qrochet.jl v0.1.2); the citation above each declaration names it.  There is
 * no FFI in the reference -- its boundary is Julia method dispatch -- so a
 * Julia shim `ccall`s these functions (a Python ctypes mirror ships beside the
 * library).  Conventions:
 *
 *  - Sites and bonds are 1-based like the reference: bond b joins sites b and b+1.
 *  - Complex numbers are interleaved (re, im) doubles, i.e. Julia ComplexF64 /
 *    C `double _Complex` / numpy complex128.
 *  - Host site arrays use the reference's default index order (o, l, r),
 *    column-major (Julia memory order): site 1 is (o, r), site n is (o, l)
 *    (`src/Ansatz/Chain.jl:33,64-100`).
 *  - Gates are column-major (out_1.., in_1..) with lanes ascending, i.e.
 *    G[out_l, out_r, in_l, in_r] for a 2-site gate (`src/Ansatz/Dense.jl:21-34`).
 *  - The library copies host data in and out; it never keeps host pointers.
 *  - Every call returns a qh_status; qh_last_error() gives the message.
 *  - One context owns one HIP stream; calls on one context are serialised
 *    in stream order.  Distinct contexts may be used from distinct threads.
 */
#ifndef QROCHET_HIP_H
#define QROCHET_HIP_H

#include <math.h>
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qh_context qh_context;
typedef struct qh_mps qh_mps;
typedef struct qh_mpo qh_mpo;

typedef enum {
    QH_OK = 0,
    QH_ERR_ARGUMENT = -1,        /* Julia ArgumentError */
    QH_ERR_MISSING_SCHMIDT = -2, /* MissingSchmidtCoefficientsException (src/Ansatz.jl:91-99) */
    QH_ERR_DIMENSION = -3,       /* DimensionMismatch / shape errors */
    QH_ERR_ASSERTION = -4,       /* AssertionError (sites / ndims checks) */
    QH_ERR_OOM = -5,             /* device allocation failed */
    QH_ERR_HIP = -6,             /* HIP runtime error */
    QH_ERR_NOT_CONVERGED = -7,   /* Jacobi SVD hit its sweep limit */
    QH_ERR_ADMISSION = -8,       /* serving: request footprint above the budget */
    QH_ERR_NUMERIC = -9          /* NaN/Inf detected in a decomposition */
} qh_status;

/* evolve flags */
#define QH_CANONICAL 1u    /* iscanonical=true  (Vidal form, Chain.jl:669-722) */
#define QH_RENORMALIZE 2u  /* renormalize=true  (Chain.jl:651-657) */
/* contraction precision (the SVD and truncation decision stay FP64) */
#define QH_PREC_FP64 0u
#define QH_PREC_FP32 16u   /* θ build on FP32 MFMA, FP32 accumulation */
#define QH_PREC_BF16X3 32u /* θ build as split-bf16 (3 products) MFMA, FP32 accumulation */
#define QH_PREC_MASK 48u

/* directions / methods of canonize_site! (Chain.jl:339-376) */
#define QH_LEFT 0
#define QH_RIGHT 1
#define QH_METHOD_QR 0
#define QH_METHOD_SVD 1

/* "nothing" for threshold (NaN) / maxdim (negative) */
#define QH_NO_THRESHOLD NAN
#define QH_NO_MAXDIM (-1)

/* ---- context ---------------------------------------------------------- */
/* No reference counterpart (the reference has no device state). */
int qh_ctx_create(int device, qh_context** out);
int qh_ctx_destroy(qh_context* ctx);
int qh_ctx_synchronize(qh_context* ctx);
const char* qh_last_error(const qh_context* ctx);
/* Kernel-class timing with HIP events on the context stream (0 = off). */
int qh_ctx_set_timing(qh_context* ctx, int enable);
/* Accumulated timing since the last reset: total ms and launch count of the
 * named kernel class ("jacobi_round", "theta", "gemm", "writeback", ...). */
int qh_ctx_get_timing(qh_context* ctx, const char* kernel_class, double* total_ms, int64_t* launches,
                      double* algorithmic_flops, double* algorithmic_bytes);
int qh_ctx_reset_timing(qh_context* ctx);
/* Jacobi tuning: block size (8, 16, 32) and sweep limit; 0 keeps the current setting
 * (by default automatic: 16, and 32 for batches with a matrix of >= 2048 columns).  Matrices
 * still rotating at the limit get a second window of as many sweeps before the call
 * returns QH_ERR_NOT_CONVERGED. */
int qh_ctx_set_jacobi(qh_context* ctx, int block, int max_sweeps);
/* Householder QR passes that precondition the Jacobi SVD of matrices with more
 * than one block pair (0..4, default 3): X = Q1 R1, R_{j-1}^H = Q_j R_j, Jacobi on
 * R_n^H.  Changes the rounding of the results only, not their meaning (no
 * reference counterpart: a tuning knob of this SVD). */
int qh_ctx_set_qr_passes(qh_context* ctx, int passes);
/* HIP graphs for the Jacobi SVD (no reference counterpart): with enable = 1 (the
 * default) the rounds of a sweep are captured once per batched SVD and replayed as
 * one graph launch per sweep; 0 launches them one by one.  The results are the same. */
int qh_ctx_set_graphs(qh_context* ctx, int enable);
/* Column pivoting of the SVD's preconditioner (no reference counterpart): with
 * enable = 1 (the default) the first QR pass factors the matrix with its columns
 * ordered by decreasing norm, X P = Q1 R1, which grades R1 and saves Jacobi sweeps;
 * 0 factors X as it is.  Changes the rounding of the results only. */
int qh_ctx_set_pivoting(qh_context* ctx, int enable);
/* Schedule tuner of the batched SVD (no reference counterpart; the project's
 * autotuner): for each size class q = 64, 128, ... <= max_q (at most 4096) it times
 * the batched SVD of `batch` synthetic q × q matrices (flat and decaying spectra;
 * fewer above q = 512) under the context's current schedule and every candidate --
 * Jacobi block width 8 or 16 and 1..4 QR preconditioning passes -- and keeps the
 * fastest per class for this context if it beats the current schedule by at least 3 %
 * (else the current schedule stays): a matrix then gets the passes of the smallest class >= its column count,
 * and a batch the block width of its largest member's class.  report (may be NULL):
 * 5 doubles per class (q, block, passes, best ms, ms of the previous setting), at most
 * 8 classes; nclasses (may be NULL) receives their number.  qh_ctx_set_jacobi and
 * qh_ctx_set_qr_passes override the tuned values; qh_ctx_autotune_reset drops them. */
int qh_ctx_autotune(qh_context* ctx, int max_q, int batch, double* report, int* nclasses);
int qh_ctx_autotune_reset(qh_context* ctx);
/* Device memory of the context's GPU (no reference counterpart: the request loop
 * sizes its admission budget from total_bytes). */
int qh_ctx_mem_info(qh_context* ctx, size_t* free_bytes, size_t* total_bytes);

/* ---- MPS objects -------------------------------------------------------- */
/* Replaces `MPS(arrays)` / `Chain(::State, ::Open, arrays)` (src/Ansatz/Chain.jl:18,64-100)
 * followed by `adapt` to the device (ext/QrochetAdaptExt.jl:9).
 * bond_dims: n-1 entries.  arrays[i]: (o,l,r) column-major, boundary 2-D.
 * lambdas: NULL, or n-1 pointers (NULL where the bond has no Λ vector). */
int qh_mps_create(qh_context* ctx, int n, int d, const int* bond_dims, const double* const* arrays,
                  const double* const* lambdas, qh_mps** out);
/* `copy`/`deepcopy` (src/Ansatz/Chain.jl:11): independent device copy. */
/* Replaces `pMPS(arrays)` = `Chain(State(), Periodic(), arrays)` (src/Ansatz/Chain.jl:19,36-62).
 * Every array is (o, l, r) column-major (3-D); bond_dims has n entries, bond i joins
 * sites i and i+1 and bond n joins site n to site 1.  Periodic chains support
 * evolve (bonds 1..n-1, as the reference: `Gate lanes must be contiguous`),
 * truncate, contract_between, the canonicality tests, expect, overlap and norm
 * (trace over the wrap bond, cost O(n·W²·χ³) for wrap dimension W); the methods the
 * reference defines for `Open` only return QH_ERR_ARGUMENT. */
int qh_pmps_create(qh_context* ctx, int n, int d, const int* bond_dims, const double* const* arrays, qh_mps** out);
/* `rand(Chain, Open, State; n, χ, p, eltype)` (src/Ansatz/Chain.jl:223-256): a random
 * normalised open MPS with bond b of dimension min(χ, p^b, p^(n-b)), right-canonical
 * from site 2 (orthonormalised rows of uniform [0, 1) entries), site 1 an isometry
 * divided by sqrt(p).  complex_entries: 0 for eltype Float64, 1 for ComplexF64.  The
 * entries come from a counter-based hash of (seed, site, entry): the same seed gives
 * the same chain; Julia's RNG stream is not reproduced. */
int qh_mps_rand(qh_context* ctx, int n, int chi, int p, int complex_entries, unsigned long long seed, qh_mps** out);
/* periodic: 1 for a periodic chain, 0 for an open one; wrap_dim: the bond between
 * site n and site 1 (1 for an open chain) */
int qh_mps_is_periodic(const qh_mps* mps, int* periodic, int* wrap_dim);
int qh_mps_clone(qh_context* ctx, const qh_mps* src, qh_mps** out);
int qh_mps_free(qh_mps* mps);
int qh_mps_nsites(const qh_mps* mps, int* n, int* d);
/* Bond dimensions (n-1) and Λ presence flags (n-1, may be NULL). */
int qh_mps_bond_dims(qh_mps* mps, int* bond_dims, int* has_lambda);
/* Download of one site tensor (o,l,r) column-major -- the `Chain(::HIPChain)` direction. */
int qh_mps_get_site(qh_mps* mps, int site, double* host_array);
/* Λ vector of a bond; returns QH_ERR_MISSING_SCHMIDT if the bond has none
 * (`tensors(ψ; between)`, src/Ansatz.jl:78-89). */
int qh_mps_get_lambda(qh_mps* mps, int bond, double* host_lambda);
/* Device footprint in bytes (site tensors + Λ, current capacity). */
int qh_mps_footprint(const qh_mps* mps, size_t* bytes);

/* ---- gate application --------------------------------------------------- */
/* `evolve_1site!` (src/Ansatz/Chain.jl:586-603): gate is d*d, (out, in) column-major. */
int qh_evolve1(qh_context* ctx, qh_mps* mps, int site, const double* gate);
/* `evolve!(ψ, gate; threshold, maxdim, iscanonical, renormalize)` for a 2-site gate
 * (src/Ansatz/Chain.jl:543-584 -> evolve_2site! :606-661, contract_2sitewf! :669-685,
 * unpack_2sitewf! :693-722, truncate! :390-422).  gate: d^4 complex, (out_l,out_r,in_l,in_r).
 * threshold NaN -> nothing; maxdim < 0 -> nothing.  kept / discarded may be NULL; if
 * given, the call waits for the result: kept = new bond dimension, discarded =
 * sqrt(sum of squared dropped singular values). */
int qh_evolve2(qh_context* ctx, qh_mps* mps, int bond, const double* gate, double threshold, int maxdim,
               unsigned flags, int* kept, double* discarded);
/* A brickwork layer: `count` independent 2-site updates, each on its own MPS or on
 * disjoint bonds of a shared one, in one pipeline of batched launches (no
 * reference counterpart as one call: it is `count` calls of evolve! fused).
 * gates: count*d^4 complex.  thresholds / maxdims: per-update truncation (count
 * entries each, NaN / negative = nothing), or NULL for "nothing" on every update.
 * kept/discarded: count entries or NULL. */
int qh_evolve2_batch(qh_context* ctx, int count, qh_mps* const* mps, const int* bonds, const double* gates,
                     const double* thresholds, const int* maxdims, unsigned flags, int* kept, double* discarded);

/* ---- truncation & canonical forms --------------------------------------- */
/* `truncate!(ψ, bond; threshold, maxdim)` (src/Ansatz/Chain.jl:390-422). */
int qh_truncate(qh_context* ctx, qh_mps* mps, int bond, double threshold, int maxdim);
/* `contract!(ψ, :between, site1, site2; direction, delete_Λ)` (src/Ansatz/Chain.jl:309-333).
 * direction QH_LEFT: into site1, QH_RIGHT: into site2. */
int qh_contract_between(qh_context* ctx, qh_mps* mps, int site1, int site2, int direction, int delete_lambda);
/* `canonize_site!(ψ, site; direction, method)` (src/Ansatz/Chain.jl:339-376). */
int qh_canonize_site(qh_context* ctx, qh_mps* mps, int site, int direction, int method);
/* `canonize!(ψ)` (src/Ansatz/Chain.jl:469-497): Vidal form. */
int qh_canonize(qh_context* ctx, qh_mps* mps);
/* `mixed_canonize!(ψ, center)` (src/Ansatz/Chain.jl:509-524). */
int qh_mixed_canonize(qh_context* ctx, qh_mps* mps, int center);
/* `normalize!(ψ, root)` (src/Ansatz/Chain.jl:532-536). */
int qh_normalize(qh_context* ctx, qh_mps* mps, int root);
/* `isleftcanonical` / `isrightcanonical` (src/Ansatz/Chain.jl:424-458):
 * result = 1 if ||A^H A - I||_F <= atol.  dev (may be NULL) receives ||A^H A - I||_F. */
int qh_is_left_canonical(qh_context* ctx, qh_mps* mps, int site, double atol, int* result, double* dev);
int qh_is_right_canonical(qh_context* ctx, qh_mps* mps, int site, double atol, int* result, double* dev);

/* ---- contractions --------------------------------------------------------- */
/* One observable: nsites = 1 (op d*d) or 2 (op d^4, on sites (site, site+1)),
 * column-major (out.., in..) like a gate. */
typedef struct {
    int nsites;
    int site;
    const double* op;
} qh_observable;
/* `expect(ψ, observables)` (src/Ansatz/Chain.jl:724-735): ⟨ψ|O_k…O_1|ψ⟩, unnormalised,
 * one complex scalar for the whole list (applied in list order). */
int qh_expect(qh_context* ctx, qh_mps* mps, int nobs, const qh_observable* obs, double* out);
/* Many independent expectations from one pair of environment sweeps: term t is
 * the product of obs[offsets[t] .. offsets[t+1]) (no reference counterpart as one
 * call: it is nterms calls of expect).  out: nterms complex. */
int qh_expect_many(qh_context* ctx, qh_mps* mps, int nterms, const int* offsets, const qh_observable* obs,
                   double* out);
/* `overlap(a, b)` (src/Ansatz/Chain.jl:737-748): Σ a·conj(b) = ⟨b|a⟩. */
int qh_overlap(qh_context* ctx, qh_mps* a, qh_mps* b, double* out);
/* `norm(ψ)` (src/Ansatz.jl:101-109). */
int qh_norm(qh_context* ctx, qh_mps* mps, double* out);
/* qh_overlap without any device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The environments live in a fixed
 * 64 MiB device array per GPU, set up when a context is created; open chains whose
 * sweep needs more (2 chi_b*chi_a + chi_b*d*chi_a complex words at the widest
 * bond, about chi <= 1024 at d = 2) get QH_ERR_ARGUMENT.  Calls that overlap in time
 * on one GPU share that array and race: the caller serialises them. */
int qh_overlap_noalloc(qh_context* ctx, qh_mps* a, qh_mps* b, double* out);

/* ---- MPO × MPS ------------------------------------------------------------ */
/* Replaces `MPO(arrays)` = `Chain(Operator(), Open(), arrays)` (src/Ansatz/Chain.jl:20,133-172)
 * followed by `adapt` to the device.  arrays[i]: (o, i, l, r) column-major (the
 * reference's default order, Chain.jl:34), boundary tensors 3-D: (o, i, r) for the
 * first site, (o, i, l) for the last.  bond_dims: n-1 entries. */
int qh_mpo_create(qh_context* ctx, int n, int d, const int* bond_dims, const double* const* arrays, qh_mpo** out);
/* Replaces `pMPO(arrays)` = `Chain(Operator(), Periodic(), arrays)` (Chain.jl:21,102-131):
 * every array (o, i, l, r) 4-D, bond_dims n entries (bond n joins site n to site 1).
 * A periodic MPO applies (exactly) to a periodic MPS; fitting is open-only. */
int qh_pmpo_create(qh_context* ctx, int n, int d, const int* bond_dims, const double* const* arrays, qh_mpo** out);
int qh_mpo_free(qh_mpo* mpo);
/* H|ψ⟩ exactly: a new MPS with bond dimensions χ_b·D_b (Λ vectors of ψ absorbed).
 * The reference has no apply routine: it would `merge` the MPO and MPS networks
 * (src/Quantum.jl:330-348) and contract them; this is the same state as a Chain. */
int qh_mpo_apply(qh_context* ctx, const qh_mpo* mpo, const qh_mps* psi, qh_mps** out);
/* Variational compression (no reference counterpart): φ <- argmin ‖φ - H ψ‖ over MPS
 * with φ's bond dimensions, by `sweeps` left-right-left one-site fitting sweeps.  φ
 * (distinct from ψ) is the initial guess; its Λ vectors are absorbed and on return
 * it is right-orthonormal from site 2 with the orthogonality centre on site 1.
 * flags: QH_PREC_* for the environment contractions (FP64 vector ALU, FP32 or bf16x3
 * on MFMA); orthonormalisation stays FP64.  overlap: 2 doubles, ⟨φ|Hψ⟩ at the end. */
int qh_mpo_fit(qh_context* ctx, const qh_mpo* mpo, const qh_mps* psi, qh_mps* phi, int sweeps, unsigned flags,
               double* overlap);

#ifdef __cplusplus
}
#endif
#endif
