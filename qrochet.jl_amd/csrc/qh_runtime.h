// Host runtime of libqrochet_hip.so: context (stream, descriptor arena, event
// timing), MPS objects, and the batched building blocks the C ABI composes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <new>
#include <unordered_map>
#include <unordered_set>
#include <stdexcept>
#include <string>
#include <vector>

#include "qh_kernels.h"
#include "qrochet_hip.h"

namespace qh {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what);
#define QH_HIP(x) ::qh::hip_check((x), #x)

struct TimerClass {
    double ms = 0.0;
    int64_t launches = 0;
    double flops = 0.0;
    double bytes = 0.0;
};

}  // namespace qh

struct qh_context {
    int device = 0;
    int num_cus = 256;  // compute units of the device (queried at creation)
    hipStream_t stream = nullptr;
    std::string err;
    // descriptor arena: chunks of pinned host staging + device mirror.  Within one
    // API call uploads only append (a full chunk opens a new one), so a descriptor
    // stays valid for the rest of the call; chunks are recycled at the entry of
    // the next call, after the stream has drained.
    struct ArenaChunk {
        char* h = nullptr;
        char* d = nullptr;
        size_t cap = 0;
    };
    std::vector<ArenaChunk> arena;
    size_t arena_off = 0;            // offset in arena.back()
    size_t arena_first = 4u << 20;   // capacity of the first chunk
    // small pinned readback area
    char* h_read = nullptr;
    size_t read_cap = 0;
    // kernel-class timing
    bool timing = false;
    std::map<std::string, qh::TimerClass> timers;
    struct Pending {
        std::string cls;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    // Jacobi parameters
    // Jacobi block width: 8, 16 or 32 as set, or 0 for the automatic choice (16, and
    // 32 for batches with a matrix of 2048 or more columns: measured on MI355X, a
    // batch of 24 SVDs of 2048 × 2048 took 2.15 s with width 32 and 2.39 s with 16,
    // while width 16 wins at every smaller size)
    int jac_block = 0;
    int jac_max_sweeps = 40;
    int jac_sort_sweeps = 2;  // sweeps that order eigenpairs by size (then plain cyclic)
    // QR passes preconditioning SVDs with more than one block pair (0..4): X = Q1 R1,
    // R_{j-1}^H = Q_j R_j, Jacobi on R_n^H
    int jac_precondition = 3;
    int jac_max_inner = 1;    // sweep cap of the block-pair eigensolver (matrices with > 1 block pair)
    bool use_graphs = true;   // a Jacobi sweep's rounds captured once and replayed as a HIP graph
    bool jac_tail = true;     // the last sweeps as whole-sweep checks plus launches of the flagged pairs
    bool jac_defer_v = true;  // block width 16: V_IJ Q of large matrices in the next launch's V workgroups
    int jac_tail_switch = 700;  // per mille of an item's block pairs: at most that many rotated -> tail
    bool pivot = true;        // first QR pass on the columns ordered by decreasing norm
    // set by qh_ctx_autotune: QR passes per size class (columns up to the key) and the
    // Jacobi block width per largest class of a batch; empty = the settings above
    std::map<int, int> tuned_passes;
    std::map<int, int> tuned_block;
    // device buffers allocated by the entry point running now and not yet owned by
    // an MPS: freed if that call fails, so an error leaks nothing
    std::unordered_set<void*> call_allocs;
    // Device memory pool.  Freeing through the runtime costs 0.1-0.3 ms of host
    // time per buffer (stream-ordered free of a >= 1 MB block), more than many
    // kernels of a layer, so the context recycles its own blocks: sizes round up
    // to classes (4 per power of two), a freed block goes to its class's free
    // list and is reused by a later allocation.  All work of a context is on one
    // stream, so a block reused by a later launch is ordered after every earlier
    // use.  Blocks <= 1 MB are carved from 8 MB slabs.
    struct Pool {
        std::unordered_map<size_t, std::vector<void*>> free;  // class size -> free blocks
        std::unordered_map<void*, size_t> size_of;            // block -> class size
        std::unordered_set<void*> live;                       // blocks handed out
        std::vector<void*> slabs;                             // carved slabs (small classes)
        std::vector<void*> big;                               // blocks allocated one by one
        size_t reserved = 0;                                  // bytes obtained from the runtime
    } pool;
};

struct qh_mps {
    qh_context* ctx = nullptr;
    int n = 0, d = 0;
    std::vector<int> chi;            // n+1 entries: chi[b] bond b; open: chi[0] = chi[n] = 1,
                                     // periodic: chi[0] = chi[n] = the wrap bond (site n -> site 1)
    bool periodic = false;
    std::vector<qh::cplx*> site;     // n
    std::vector<double*> lam;        // n-1 (bond b -> lam[b-1]); null when absent
};

namespace qh {

// ---- context helpers ----
void ctx_sync(qh_context* c);
void* dalloc_bytes(qh_context* c, size_t bytes);
template <class T>
T* dalloc(qh_context* c, size_t n) {
    return static_cast<T*>(dalloc_bytes(c, n * sizeof(T)));
}
void dfree(qh_context* c, void* p);
// return every cached block to the runtime (the stream is drained first)
void pool_trim(qh_context* c);
void pool_destroy(qh_context* c);
// `p` now belongs to an MPS and outlives the current call
void keep(qh_context* c, void* p);
// free what the failed call left behind
void release_call_allocs(qh_context* c);
// copy host bytes to the descriptor arena; the device pointer stays valid until the
// calling API entry point returns
void* arena_upload(qh_context* c, const void* src, size_t bytes);
// at API entry: drain the stream and recycle the arena (keeps the largest chunk)
void arena_recycle(qh_context* c);
void arena_free(qh_context* c);
template <class T>
T* upload(qh_context* c, const std::vector<T>& v) {
    return static_cast<T*>(arena_upload(c, v.data(), v.size() * sizeof(T)));
}
// several uploads with one host-to-device copy
void arena_upload_segments(qh_context* c, int count, const void* const* srcs, const size_t* bytes, void** outs);
template <class A, class B>
void upload2(qh_context* c, const std::vector<A>& a, const std::vector<B>& b, A*& da, B*& db) {
    const void* srcs[2] = {a.data(), b.data()};
    const size_t bytes[2] = {a.size() * sizeof(A), b.size() * sizeof(B)};
    void* outs[2];
    arena_upload_segments(c, 2, srcs, bytes, outs);
    da = static_cast<A*>(outs[0]);
    db = static_cast<B*>(outs[1]);
}
// device -> host readback (synchronises)
void readback(qh_context* c, void* host, const void* dev, size_t bytes);

// timing scope around one launch
struct Timed {
    qh_context* c;
    const char* cls;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(qh_context* c, const char* cls, double flops = 0.0, double bytes = 0.0);
    ~Timed();
};

// With `device` current, once per device: the device address of the fixed scratch
// array of qh_overlap_noalloc (throws on failure).
void noalloc_setup(int device);

// ---- batched building blocks ----
void run_theta(qh_context* c, std::vector<ThetaItem>& items, int precision = QH_PREC_FP64);
// the GEMM launch itself (no split-K)
void run_gemm_tiles(qh_context* c, std::vector<GemmItem>& items, int precision = QH_PREC_FP64);
void run_gemm(qh_context* c, std::vector<GemmItem>& items, int precision = QH_PREC_FP64);
// runs set-identity + sweeps until every item converged; throws on sweep limit.
// fro2_out (optional, one double per item) receives ||X||_F² and stays valid after.
// block: column block width (8 or 16; 0 = the context's setting)
void run_jacobi(qh_context* c, std::vector<JacItem>& items, double* fro2_out = nullptr, int block = 0);
void run_select(qh_context* c, std::vector<SelectItem>& items);
void run_gather(qh_context* c, std::vector<GatherItem>& items);

GemmItem gemm_item(const cplx* a, int opa, int lda, const cplx* b, int opb, int ldb, cplx* cc, int ldc, int m, int n,
                   int k);

// ---- MPS helpers ----
void check_mps(qh_context* c, const qh_mps* m);
size_t site_elems(const qh_mps* m, int site);  // 1-based
void replace_site(qh_mps* m, int site, cplx* buf);
void replace_lam(qh_mps* m, int bond, double* buf);

// One SVD of a logical matrix M (mr × mc) held in the caller's buffer x (as M
// if mr >= mc, else as M^H), with the factor outputs the caller needs.
struct SvdOut {
    int factor = 0;          // 0 = U_M, 1 = V_M
    int conj_transpose = 0;  // write the factor's conjugate transpose
    const double* rowscale = nullptr;
    int rowscale_mod = 1, rowscale_div = 1;
    cplx* dst = nullptr;
};
struct SvdTask {
    cplx* x = nullptr;   // M (mr × mc) if mr >= mc, else M^H (mc × mr); destroyed
    int mr = 0, mc = 0;
    int truncate = 0;
    double threshold = 1e-16;
    int maxdim = -1;
    int renormalize = 0;
    // one Newton-Schulz (polar) step on both kept factors (F <- F (3I - F^H F) / 2):
    // their orthogonality goes from what the Jacobi's tolerance and its accumulated
    // rotations leave (~1e-12 in Frobenius norm at 256 columns) to rounding level, like
    // LAPACK's factors -- for callers whose result must pass the canonical-form tests
    int refine = 0;
    std::vector<SvdOut> outs;
    double* lam_dst = nullptr;     // kept singular values (may be null)
    double* sigma_dst = nullptr;   // all sorted singular values (k), may be null
    int kept = 0;                  // result
    double discarded = 0.0;        // result
    int kmax() const;
};
// one dense rows × cols matrix (ld rows) for polar_refine
struct PolarJob {
    cplx* f;
    int rows, cols;
};
// F <- F (3I - F^H F) / 2 for every job, in place (one Newton-Schulz step toward the
// nearest isometry; zero columns stay zero)
void polar_refine(qh_context* c, const std::vector<PolarJob>& jobs);

// Runs the batch of SVD tasks (Jacobi + select + gathers) and synchronises.
void run_svd_tasks(qh_context* c, std::vector<SvdTask>& tasks);

// most QR passes of the Jacobi preconditioner (qh_context::jac_precondition)
constexpr int QR_MAX_PASSES = 4;
// the passes a matrix with q columns gets (tuned table or the context's setting)
int qr_passes_for(const qh_context* c, int q);
// the Jacobi block width of a batch whose widest matrix has qmax columns
int block_for(const qh_context* c, int qmax);
// Householder QR of every job in place (R in the upper triangle); reflectors into v (m × n),
// compact-WY triangles into t (ceil(n / QR_NB) · QR_NB² entries).
struct QrJob {
    cplx* a;
    int m, n, lda;
    cplx* v;
    cplx* t;
};
void run_qr_factor(qh_context* c, const std::vector<QrJob>& jobs);
// y (m × ncols, ld m) <- Q y for every job (Q = the factorisation's orthogonal factor)
struct QrApplyJob {
    const QrJob* qr;
    cplx* y;
    int ncols;
};
void run_qr_apply_q(qh_context* c, const std::vector<QrApplyJob>& jobs);


// ---- entry-point wrapper ----
// error text of calls made without a context (per thread)
std::string& thread_error();
// Runs one C-ABI entry point: device selection, per-call allocation tracking and
// arena recycling; exceptions become status codes, and a failed call releases
// the device buffers it allocated.
template <class F>
int api_guard(qh_context* c, F&& f) {
    try {
        if (c) {
            QH_HIP(hipSetDevice(c->device));
            c->call_allocs.clear();
            arena_recycle(c);
        }
        f();
        if (c) {
            c->err.clear();
            c->call_allocs.clear();
        }
        return QH_OK;
    } catch (const Error& e) {
        if (c) release_call_allocs(c);
        (c ? c->err : thread_error()) = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        if (c) release_call_allocs(c);
        (c ? c->err : thread_error()) = "host allocation failed";
        return QH_ERR_OOM;
    } catch (const std::exception& e) {
        if (c) release_call_allocs(c);
        (c ? c->err : thread_error()) = e.what();
        return QH_ERR_HIP;
    }
}

}  // namespace qh
