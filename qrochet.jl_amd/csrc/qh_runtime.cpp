// Host runtime: context services and the batched building blocks
// (θ build, GEMM, Jacobi SVD, select, gather) shared by every entry point.
#include "qh_runtime.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace qh {

void launch_gather_items(const GatherItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s);

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(QH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// context services
// ---------------------------------------------------------------------------
static void collect_timers(qh_context* c) {
    for (auto& p : c->pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) c->timers[p.cls].ms += ms;
        c->event_pool.push_back(p.a);
        c->event_pool.push_back(p.b);
    }
    c->pending.clear();
}

void ctx_sync(qh_context* c) {
    QH_HIP(hipStreamSynchronize(c->stream));
    c->arena_off = 0;
    if (!c->pending.empty()) collect_timers(c);
}

void* dalloc_bytes(qh_context* c, size_t bytes) {
    void* p = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMallocAsync(&p, bytes, c->stream);
    if (e != hipSuccess || p == nullptr)
        throw Error(QH_ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    return p;
}

void dfree(qh_context* c, void* p) {
    if (p) (void)hipFreeAsync(p, c->stream);
}

void* arena_upload(qh_context* c, const void* src, size_t bytes) {
    const size_t need = (bytes + 255) & ~size_t(255);
    if (c->arena_off + need > c->arena_cap) {
        ctx_sync(c);  // every earlier upload has been consumed
        if (need > c->arena_cap) {
            size_t cap = std::max(need, c->arena_cap * 2);
            if (c->h_arena) (void)hipHostFree(c->h_arena);
            if (c->d_arena) (void)hipFree(c->d_arena);
            c->h_arena = nullptr;
            c->d_arena = nullptr;
            QH_HIP(hipHostMalloc((void**)&c->h_arena, cap, hipHostMallocDefault));
            QH_HIP(hipMalloc((void**)&c->d_arena, cap));
            c->arena_cap = cap;
        }
    }
    char* h = c->h_arena + c->arena_off;
    char* d = c->d_arena + c->arena_off;
    if (bytes) {
        std::memcpy(h, src, bytes);
        QH_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream));
    }
    c->arena_off += need;
    return d;
}

void readback(qh_context* c, void* host, const void* dev, size_t bytes) {
    if (bytes == 0) return;
    if (bytes > c->read_cap) {
        if (c->h_read) (void)hipHostFree(c->h_read);
        c->h_read = nullptr;
        size_t cap = std::max(bytes, size_t(1) << 16);
        QH_HIP(hipHostMalloc((void**)&c->h_read, cap, hipHostMallocDefault));
        c->read_cap = cap;
    }
    QH_HIP(hipMemcpyAsync(c->h_read, dev, bytes, hipMemcpyDeviceToHost, c->stream));
    ctx_sync(c);
    std::memcpy(host, c->h_read, bytes);
}

static hipEvent_t get_event(qh_context* c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    QH_HIP(hipEventCreate(&e));
    return e;
}

Timed::Timed(qh_context* c_, const char* cls_, double flops, double bytes) : c(c_), cls(cls_) {
    if (!c->timing) return;
    auto& t = c->timers[cls];
    t.launches += 1;
    t.flops += flops;
    t.bytes += bytes;
    a = get_event(c);
    b = get_event(c);
    QH_HIP(hipEventRecord(a, c->stream));
}
Timed::~Timed() {
    if (!a) return;
    (void)hipEventRecord(b, c->stream);
    c->pending.push_back({cls, a, b});
}

// ---------------------------------------------------------------------------
// θ build
// ---------------------------------------------------------------------------
void run_theta(qh_context* c, std::vector<ThetaItem>& items) {
    if (items.empty()) return;
    const int d = items[0].d;
    const int te = theta_tile_extent(d);
    std::vector<int2> tiles;
    double flops = 0, bytes = 0;
    for (size_t i = 0; i < items.size(); ++i) {
        ThetaItem& t = items[i];
        t.tiles_l = (t.chil + te - 1) / te;
        t.tiles_r = (t.chir + te - 1) / te;
        for (int k = 0; k < t.tiles_l * t.tiles_r; ++k) tiles.push_back(make_int2((int)i, k));
        const double dd = d;
        flops += 8.0 * dd * dd * t.chil * (double)t.chim * t.chir + 8.0 * dd * dd * dd * dd * t.chil * (double)t.chir;
        bytes += 16.0 * (dd * t.chil * t.chim + dd * t.chim * (double)t.chir + dd * dd * t.chil * (double)t.chir);
    }
    ThetaItem* di = upload(c, items);
    int2* dt = upload(c, tiles);
    Timed tm(c, "theta", flops, bytes);
    launch_theta(di, dt, (int)tiles.size(), d, c->stream);
}

// ---------------------------------------------------------------------------
// GEMM
// ---------------------------------------------------------------------------
GemmItem gemm_item(const cplx* a, int opa, int lda, const cplx* b, int opb, int ldb, cplx* cc, int ldc, int m, int n,
                   int k) {
    GemmItem g{};
    g.a = a; g.b = b; g.c = cc;
    g.opa = opa; g.opb = opb;
    g.lda = lda; g.ldb = ldb; g.ldc = ldc;
    g.m = m; g.n = n; g.k = k;
    g.alpha = make_double2(1.0, 0.0);
    g.beta = make_double2(0.0, 0.0);
    g.dk = nullptr; g.dk_mod = 1; g.dk_div = 1;
    g.rowscale = nullptr; g.colscale = nullptr;
    return g;
}

void run_gemm(qh_context* c, std::vector<GemmItem>& items) {
    if (items.empty()) return;
    std::vector<int2> tiles;
    double flops = 0, bytes = 0;
    for (size_t i = 0; i < items.size(); ++i) {
        GemmItem& g = items[i];
        g.tiles_m = (g.m + GEMM_TM - 1) / GEMM_TM;
        g.tiles_n = (g.n + GEMM_TN - 1) / GEMM_TN;
        if (g.dk_mod <= 0) g.dk_mod = 1;
        if (g.dk_div <= 0) g.dk_div = 1;
        for (int k = 0; k < g.tiles_m * g.tiles_n; ++k) tiles.push_back(make_int2((int)i, k));
        flops += 8.0 * g.m * (double)g.n * g.k;
        bytes += 16.0 * ((double)g.m * g.k + (double)g.k * g.n + (double)g.m * g.n);
    }
    if (tiles.empty()) return;
    GemmItem* di = upload(c, items);
    int2* dt = upload(c, tiles);
    Timed tm(c, "gemm", flops, bytes);
    launch_gemm(di, dt, (int)tiles.size(), c->stream);
}

// ---------------------------------------------------------------------------
// Jacobi SVD
// ---------------------------------------------------------------------------
void run_jacobi(qh_context* c, std::vector<JacItem>& items) {
    if (items.empty()) return;
    const int block = c->jac_block;
    const int bc = 2 * block;
    std::vector<int2> id_tiles, tiles;
    int rmax = 1;
    int* counters = dalloc<int>(c, 2 * items.size() + 1);
    QH_HIP(hipMemsetAsync(counters, 0, (2 * items.size() + 1) * sizeof(int), c->stream));
    double* fro2 = dalloc<double>(c, items.size());
    QH_HIP(hipMemsetAsync(fro2, 0, items.size() * sizeof(double), c->stream));
    std::vector<int2> frob_tiles;
    for (size_t i = 0; i < items.size(); ++i) {
        JacItem& it = items[i];
        int nb = (it.q + block - 1) / block;
        if (nb < 2) nb = 2;
        if (nb & 1) ++nb;
        it.nb = nb;
        it.rot = counters + 2 * i;
        it.conv = counters + 2 * i + 1;
        it.fro2 = fro2 + i;
        const size_t tot = (size_t)it.p * it.q;
        for (size_t k = 0; k * FROB_CHUNK < tot; ++k) frob_tiles.push_back(make_int2((int)i, (int)k));
        rmax = std::max(rmax, nb - 1);
        for (int k = 0; k < it.q; ++k) id_tiles.push_back(make_int2((int)i, k));
        for (int k = 0; k < nb / 2; ++k) tiles.push_back(make_int2((int)i, k));
    }
    int* remaining = counters + 2 * items.size();
    JacItem* di = upload(c, items);
    int2* dti = upload(c, id_tiles);
    int2* dt = upload(c, tiles);
    int2* dtf = upload(c, frob_tiles);
    launch_set_identity_items(di, dti, (int)id_tiles.size(), c->stream);
    launch_frob2(di, dtf, (int)frob_tiles.size(), c->stream);
    // Relative orthogonality target |x_i^H x_j| <= tol ||x_i|| ||x_j||.  The Gram
    // entries carry O(sqrt(p)·ε) rounding noise, so the target sits a factor 16
    // above it; otherwise noise-triggered rotations never stop.
    int pmax = 1;
    for (const JacItem& it : items) pmax = std::max(pmax, it.p);
    const double tol = 16.0 * std::sqrt((double)std::max(pmax, 4)) * DBL_EPSILON;
    const double floor_rel = 4.0 * std::sqrt((double)std::max(pmax, 4)) * DBL_EPSILON;
    const double floor_rel2 = floor_rel * floor_rel;
    // algorithmic work per checked / rotated block pair (for the timing report)
    std::vector<double> gram_f(items.size()), upd_f(items.size()), gram_b(items.size()), upd_b(items.size());
    for (size_t i = 0; i < items.size(); ++i) {
        const double p = items[i].p, q = items[i].q;
        gram_f[i] = 4.0 * bc * (bc + 1) * p;
        upd_f[i] = 8.0 * bc * bc * (p + q);
        gram_b[i] = 16.0 * bc * p;
        upd_b[i] = 32.0 * bc * (p + q);
    }
    std::vector<char> converged(items.size(), 0);
    std::vector<int> hrot(2 * items.size());
    bool done = false;
    for (int sweep = 0; sweep < c->jac_max_sweeps; ++sweep) {
        // read the per-item rotation counters after the window to account work
        for (int r = 0; r < rmax; ++r) {
            const int g = sweep * rmax + r;
            Timed tm(c, "jacobi_round");
            static const int kdbg = std::getenv("QH_DEBUG_JACOBI_KERNEL") ? 1 : 0;
            launch_jacobi_round(di, dt, (int)tiles.size(), g, block, tol, floor_rel2, sweep < c->jac_sort_sweeps ? 1 : 0,
                                kdbg, c->stream);
        }
        static const bool dbg = std::getenv("QH_DEBUG_JACOBI") != nullptr;
        if (dbg) {
            readback(c, hrot.data(), counters, hrot.size() * sizeof(int));
            std::fprintf(stderr, "[jacobi] sweep %d rmax %d rot:", sweep, rmax);
            for (size_t i = 0; i < items.size(); ++i)
                std::fprintf(stderr, " %d/%d(%dx%d,nb%d)", hrot[2 * i], hrot[2 * i + 1], items[i].p, items[i].q,
                             items[i].nb);
            std::fprintf(stderr, "\n");
        }
        if (c->timing) {
            readback(c, hrot.data(), counters, hrot.size() * sizeof(int));
            auto& t = c->timers["jacobi_round"];
            for (size_t i = 0; i < items.size(); ++i) {
                if (converged[i]) continue;
                const double checked = (double)(items[i].nb / 2) * rmax;
                t.flops += checked * gram_f[i] + hrot[2 * i] * upd_f[i];
                t.bytes += checked * gram_b[i] + hrot[2 * i] * upd_b[i];
                if (hrot[2 * i] == 0) converged[i] = 1;
            }
        }
        launch_conv_update(di, (int)items.size(), remaining, c->stream);
        c->timers["jacobi_sweep"].launches += 1;
        int rem = 0;
        readback(c, &rem, remaining, sizeof(int));
        if (rem == 0) {
            done = true;
            break;
        }
    }
    if (!done) {
        std::vector<int> h(2 * items.size());
        readback(c, h.data(), counters, h.size() * sizeof(int));
        std::string msg = "Jacobi SVD did not converge within " + std::to_string(c->jac_max_sweeps) + " sweeps:";
        if (const char* dump = std::getenv("QH_DEBUG_JACOBI_DUMP")) {
            for (size_t i = 0; i < items.size(); ++i)
                if (!h[2 * i + 1]) {
                    std::vector<cplx> hx((size_t)items[i].p * items[i].q);
                    readback(c, hx.data(), items[i].x, hx.size() * sizeof(cplx));
                    if (FILE* f = std::fopen(dump, "wb")) {
                        int dims[2] = {items[i].p, items[i].q};
                        std::fwrite(dims, sizeof(int), 2, f);
                        std::fwrite(hx.data(), sizeof(cplx), hx.size(), f);
                        std::fclose(f);
                    }
                    break;
                }
        }
        int shown = 0;
        for (size_t i = 0; i < items.size() && shown < 4; ++i)
            if (!h[2 * i + 1]) {
                msg += " item " + std::to_string(i) + " (" + std::to_string(items[i].p) + "x" +
                       std::to_string(items[i].q) + ")";
                ++shown;
            }
        dfree(c, counters);
        dfree(c, fro2);
        throw Error(QH_ERR_NOT_CONVERGED, msg);
    }
    dfree(c, counters);
    dfree(c, fro2);
}

// ---------------------------------------------------------------------------
// select + gather
// ---------------------------------------------------------------------------
void run_select(qh_context* c, std::vector<SelectItem>& items) {
    if (items.empty()) return;
    std::vector<int2> tiles;
    for (size_t i = 0; i < items.size(); ++i) {
        if (items[i].q > SELECT_MAX_Q) throw Error(QH_ERR_DIMENSION, "SVD with more than 4096 columns");
        for (int k = 0; k < (items[i].q + 3) / 4; ++k) tiles.push_back(make_int2((int)i, k));
    }
    SelectItem* di = upload(c, items);
    int2* dt = upload(c, tiles);
    Timed tm(c, "select");
    launch_select_items(di, dt, (int)tiles.size(), (int)items.size(), c->stream);
}

void run_gather(qh_context* c, std::vector<GatherItem>& items) {
    if (items.empty()) return;
    std::vector<int2> tiles;
    double bytes = 0;
    for (size_t i = 0; i < items.size(); ++i) {
        const int ntc = (items[i].rows + 31) / 32, ntj = (items[i].max_cols + 31) / 32;
        for (int k = 0; k < ntc * ntj; ++k) tiles.push_back(make_int2((int)i, k));
        bytes += 32.0 * items[i].rows * (double)items[i].max_cols;
    }
    if (tiles.empty()) return;
    GatherItem* di = upload(c, items);
    int2* dt = upload(c, tiles);
    Timed tm(c, "gather", 0.0, bytes);
    launch_gather_items(di, dt, (int)tiles.size(), c->stream);
}

int SvdTask::kmax() const {
    const int k = std::min(mr, mc);
    if (truncate && maxdim >= 0) return std::min(k, maxdim);
    return k;
}

void run_svd_tasks(qh_context* c, std::vector<SvdTask>& tasks) {
    if (tasks.empty()) return;
    // workspace: V (q×q), sigma (q), perm (q), kept (int), discarded (double)
    size_t bytes = 0;
    std::vector<size_t> off_v(tasks.size()), off_s(tasks.size()), off_p(tasks.size()), off_l(tasks.size());
    auto take = [&](size_t b) {
        size_t o = bytes;
        bytes += (b + 255) & ~size_t(255);
        return o;
    };
    for (size_t i = 0; i < tasks.size(); ++i) {
        const size_t q = std::min(tasks[i].mr, tasks[i].mc);
        off_v[i] = take(q * q * sizeof(cplx));
        off_s[i] = take(q * sizeof(double));
        off_p[i] = take(q * sizeof(int));
        off_l[i] = take(q * sizeof(double));
    }
    const size_t off_k = take(tasks.size() * sizeof(int));
    const size_t off_d = take(tasks.size() * sizeof(double));
    char* ws = static_cast<char*>(dalloc_bytes(c, bytes));
    int* d_kept = reinterpret_cast<int*>(ws + off_k);
    double* d_disc = reinterpret_cast<double*>(ws + off_d);
    std::vector<JacItem> jac(tasks.size());
    std::vector<SelectItem> sel(tasks.size());
    for (size_t i = 0; i < tasks.size(); ++i) {
        const SvdTask& t = tasks[i];
        const int p = std::max(t.mr, t.mc), q = std::min(t.mr, t.mc);
        jac[i] = JacItem{};
        jac[i].x = t.x;
        jac[i].v = reinterpret_cast<cplx*>(ws + off_v[i]);
        jac[i].p = p;
        jac[i].q = q;
        SelectItem& s = sel[i];
        s = SelectItem{};
        s.x = t.x;
        s.p = p;
        s.q = q;
        s.sigma_sorted = reinterpret_cast<double*>(ws + off_s[i]);
        s.perm = reinterpret_cast<int*>(ws + off_p[i]);
        s.lam_out = t.lam_dst ? t.lam_dst : reinterpret_cast<double*>(ws + off_l[i]);
        s.kept = d_kept + i;
        s.discarded = d_disc + i;
        s.threshold = t.threshold;
        s.maxdim = t.maxdim;
        s.truncate = t.truncate;
        s.renormalize = t.renormalize;
    }
    run_jacobi(c, jac);
    run_select(c, sel);
    std::vector<GatherItem> gat;
    for (size_t i = 0; i < tasks.size(); ++i) {
        const SvdTask& t = tasks[i];
        const bool xh = t.mr < t.mc;  // X holds M^H
        const int p = std::max(t.mr, t.mc), q = std::min(t.mr, t.mc);
        for (const SvdOut& o : t.outs) {
            GatherItem g{};
            // factor rows: U_M has mr rows, V_M has mc rows
            // source: columns of X (normalised by σ) or of the Jacobi V
            const bool from_x = (o.factor == 0) != xh;
            g.src = from_x ? t.x : jac[i].v;
            g.ld_src = from_x ? p : q;
            g.rows = (o.factor == 0) ? t.mr : t.mc;
            g.perm = sel[i].perm;
            g.sigma_sorted = sel[i].sigma_sorted;
            g.divide_sigma = from_x ? 1 : 0;
            g.kept = sel[i].kept;
            g.rowscale = o.rowscale;
            g.rowscale_mod = o.rowscale_mod > 0 ? o.rowscale_mod : 1;
            g.rowscale_div = o.rowscale_div > 0 ? o.rowscale_div : 1;
            g.out = o.dst;
            g.conj_transpose = o.conj_transpose;
            g.max_cols = t.kmax();
            gat.push_back(g);
        }
    }
    run_gather(c, gat);
    std::vector<int> kept(tasks.size());
    std::vector<double> disc(tasks.size());
    readback(c, kept.data(), d_kept, kept.size() * sizeof(int));
    readback(c, disc.data(), d_disc, disc.size() * sizeof(double));
    for (size_t i = 0; i < tasks.size(); ++i) {
        tasks[i].kept = kept[i];
        tasks[i].discarded = disc[i];
    }
    dfree(c, ws);
}

// ---------------------------------------------------------------------------
// MPS helpers
// ---------------------------------------------------------------------------
void check_mps(qh_context* c, const qh_mps* m) {
    if (!m) throw Error(QH_ERR_ARGUMENT, "null MPS handle");
    if (m->ctx != c) throw Error(QH_ERR_ARGUMENT, "MPS belongs to another context");
}

size_t site_elems(const qh_mps* m, int site) {
    return (size_t)m->chi[site - 1] * m->d * m->chi[site];
}

void replace_site(qh_mps* m, int site, cplx* buf) {
    dfree(m->ctx, m->site[site - 1]);
    m->site[site - 1] = buf;
}

void replace_lam(qh_mps* m, int bond, double* buf) {
    dfree(m->ctx, m->lam[bond - 1]);
    m->lam[bond - 1] = buf;
}

}  // namespace qh
