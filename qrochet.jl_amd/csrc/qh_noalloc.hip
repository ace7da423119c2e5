// `overlap` for callers that cannot allocate device memory: the environment sweep
// of qh_overlap (same GEMMs, same kernel, same order) with its intermediate
// environments in a fixed-size device array instead of the context's pool.
//
// The array lives for the life of the process, one copy per GPU the library runs
// on.  Its device address is looked up once per device, with that device current,
// when a context is created (hipGetSymbolAddress); host code uses only that
// address, never the array's name.  Every call on one GPU shares the array and
// leaves the last call's contents in it: calls that overlap in time on one GPU
// (two contexts, or two threads) race and must be serialised by the caller.
#include <mutex>

#include "qh_runtime.h"

namespace qh {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
void* g_scratch[NOALLOC_MAX_DEVICES] = {};

__global__ void k_set_unit(double2* e) {
    if (threadIdx.x == 0 && blockIdx.x == 0) e[0] = make_double2(1.0, 0.0);
}

size_t round256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

// With `device` current: resolve this device's copy of the scratch array (once).
void noalloc_setup(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) throw Error(QH_ERR_ARGUMENT, "device ordinal beyond 64");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        throw Error(QH_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = p;
}

static void* scratch_of(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return (device >= 0 && device < NOALLOC_MAX_DEVICES) ? g_scratch[device] : nullptr;
}

}  // namespace qh

using namespace qh;

extern "C" int qh_overlap_noalloc(qh_context* c, qh_mps* a, qh_mps* b, double* out) {
    if (!c || !out) return QH_ERR_ARGUMENT;
    try {
        QH_HIP(hipSetDevice(c->device));
        check_mps(c, a);
        check_mps(c, b);
        if (a->n != b->n || a->d != b->d) throw Error(QH_ERR_ASSERTION, "Ansatzes must have the same sites");
        if (a->periodic || b->periodic)
            throw Error(QH_ERR_ARGUMENT, "qh_overlap_noalloc contracts open chains only (use qh_overlap)");
        const int n = a->n, d = a->d;
        // E_i (chi_b[i] x chi_a[i]) ping-pongs between two buffers; T_i (chi_b[i-1] x d chi_a[i])
        size_t emax = 1, tmax = 1;
        for (int i = 0; i <= n; ++i) emax = std::max(emax, (size_t)b->chi[i] * a->chi[i]);
        for (int i = 1; i <= n; ++i) tmax = std::max(tmax, (size_t)b->chi[i - 1] * d * a->chi[i]);
        const size_t need = 2 * round256(emax * sizeof(cplx)) + round256(tmax * sizeof(cplx));
        if (need > NOALLOC_SCRATCH_BYTES)
            throw Error(QH_ERR_ARGUMENT, "qh_overlap_noalloc needs " + std::to_string(need) +
                                             " bytes of scratch, more than its " +
                                             std::to_string(NOALLOC_SCRATCH_BYTES) + " (use qh_overlap)");
        char* base = static_cast<char*>(scratch_of(c->device));
        if (!base) throw Error(QH_ERR_HIP, "scratch of this device not set up (context creation failed?)");
        cplx* E = reinterpret_cast<cplx*>(base);
        cplx* E2 = reinterpret_cast<cplx*>(base + round256(emax * sizeof(cplx)));
        cplx* T = reinterpret_cast<cplx*>(base + 2 * round256(emax * sizeof(cplx)));
        k_set_unit<<<1, 64, 0, c->stream>>>(E);
        QH_HIP(hipGetLastError());
        for (int i = 1; i <= n; ++i) {
            const int la = a->chi[i - 1], ra = a->chi[i], lb = b->chi[i - 1], rb = b->chi[i];
            // T = E diag(Λa_{i-1}) A_i   (lb x d ra)
            GemmItem g = gemm_item(E, GEMM_N, lb, a->site[i - 1], GEMM_N, la, T, lb, lb, d * ra, la);
            if (i >= 2) { g.dk = a->lam[i - 2]; g.dk_mod = la; }
            launch_gemm_arg(g, c->stream);
            // E' = (diag(Λb_{i-1}) B_i)^H T   (rb x ra), rows of B_i and T are (l, o)
            GemmItem h = gemm_item(b->site[i - 1], GEMM_C, lb * d, T, GEMM_N, lb * d, E2, rb, rb, ra, lb * d);
            if (i >= 2) { h.dk = b->lam[i - 2]; h.dk_mod = lb; }
            launch_gemm_arg(h, c->stream);
            std::swap(E, E2);
        }
        QH_HIP(hipGetLastError());
        cplx v;
        QH_HIP(hipMemcpyAsync(&v, E, sizeof(cplx), hipMemcpyDeviceToHost, c->stream));
        QH_HIP(hipStreamSynchronize(c->stream));
        out[0] = v.x;
        out[1] = v.y;
        c->err.clear();
        return QH_OK;
    } catch (const Error& e) {
        c->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        c->err = e.what();
        return QH_ERR_HIP;
    }
}
