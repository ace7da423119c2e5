// Host runtime: context services and the batched building blocks
// (θ build, GEMM, Jacobi SVD, select, gather) shared by every entry point.
#include "qh_runtime.h"

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <climits>
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

std::string& thread_error() {
    static thread_local std::string e;
    return e;
}

void ctx_sync(qh_context* c) {
    QH_HIP(hipStreamSynchronize(c->stream));
    if (!c->pending.empty()) collect_timers(c);
}

namespace {

size_t pool_class(size_t bytes) {
    if (bytes <= 4096) return (bytes + 255) & ~size_t(255);
    size_t base = 1;
    while ((base << 1) <= bytes) base <<= 1;
    const size_t step = base / 4;
    return (bytes + step - 1) / step * step;
}

constexpr size_t POOL_SLAB = size_t(8) << 20;
constexpr size_t POOL_SMALL = size_t(1) << 20;

void* pool_alloc(qh_context* c, size_t bytes) {
    auto& P = c->pool;
    const size_t cs = pool_class(bytes);
    auto& fl = P.free[cs];
    if (fl.empty()) {
        auto grab = [&](size_t n) -> void* {
            void* p = nullptr;
            if (hipMalloc(&p, n) != hipSuccess || !p) {
                (void)hipGetLastError();
                pool_trim(c);
                p = nullptr;
                if (hipMalloc(&p, n) != hipSuccess || !p) {
                    (void)hipGetLastError();
                    return nullptr;
                }
            }
            P.reserved += n;
            return p;
        };
        if (cs <= POOL_SMALL) {
            char* slab = static_cast<char*>(grab(POOL_SLAB));
            if (!slab) return nullptr;
            P.slabs.push_back(slab);
            auto& fl2 = P.free[cs];  // (pool_trim may have rehashed the map)
            for (size_t off = 0; off + cs <= POOL_SLAB; off += cs) {
                fl2.push_back(slab + off);
                P.size_of[slab + off] = cs;
            }
        } else {
            void* b = grab(cs);
            if (!b) return nullptr;
            P.big.push_back(b);
            P.size_of[b] = cs;
            P.free[cs].push_back(b);
        }
    }
    auto& fl3 = P.free[cs];
    void* p = fl3.back();
    fl3.pop_back();
    P.live.insert(p);
    return p;
}

}  // namespace

void* dalloc_bytes(qh_context* c, size_t bytes) {
    if (bytes == 0) bytes = 16;
    void* p = pool_alloc(c, bytes);
    if (!p) throw Error(QH_ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
    c->call_allocs.insert(p);
    return p;
}

void dfree(qh_context* c, void* p) {
    if (!p) return;
    c->call_allocs.erase(p);
    // only a block that is handed out returns to its free list (a second free of
    // the same pointer is ignored instead of handing the block out twice)
    if (!c->pool.live.erase(p)) return;
    auto it = c->pool.size_of.find(p);
    if (it != c->pool.size_of.end()) c->pool.free[it->second].push_back(p);
}

void pool_trim(qh_context* c) {
    // only whole big blocks can go back; slabs stay (their blocks may be in use)
    auto& P = c->pool;
    (void)hipStreamSynchronize(c->stream);
    std::unordered_set<void*> freed;
    for (auto& kv : P.free) {
        if (kv.first <= POOL_SMALL) continue;
        for (void* b : kv.second) {
            (void)hipFree(b);
            P.size_of.erase(b);
            P.reserved -= kv.first;
            freed.insert(b);
        }
        kv.second.clear();
    }
    std::vector<void*> keep_big;
    for (void* b : P.big)
        if (!freed.count(b)) keep_big.push_back(b);
    P.big.swap(keep_big);
}

void pool_destroy(qh_context* c) {
    auto& P = c->pool;
    for (void* b : P.big) (void)hipFree(b);
    for (void* s : P.slabs) (void)hipFree(s);
    P.big.clear();
    P.slabs.clear();
    P.free.clear();
    P.size_of.clear();
    P.live.clear();
    P.reserved = 0;
}

void keep(qh_context* c, void* p) {
    if (p) c->call_allocs.erase(p);
}

void release_call_allocs(qh_context* c) {
    std::vector<void*> v(c->call_allocs.begin(), c->call_allocs.end());
    for (void* p : v) dfree(c, p);
    c->call_allocs.clear();
}

void arena_upload_segments(qh_context* c, int count, const void* const* srcs, const size_t* bytes, void** outs) {
    size_t need = 0;
    for (int i = 0; i < count; ++i) need += (bytes[i] + 255) & ~size_t(255);
    if (c->arena.empty() || c->arena_off + need > c->arena.back().cap) {
        qh_context::ArenaChunk ch;
        ch.cap = std::max(need, c->arena.empty() ? c->arena_first : c->arena.back().cap * 2);
        QH_HIP(hipHostMalloc((void**)&ch.h, ch.cap, hipHostMallocDefault));
        if (hipMalloc((void**)&ch.d, ch.cap) != hipSuccess) {
            (void)hipHostFree(ch.h);
            throw Error(QH_ERR_OOM, "descriptor arena allocation failed");
        }
        c->arena.push_back(ch);
        c->arena_off = 0;
    }
    char* h = c->arena.back().h + c->arena_off;
    char* d = c->arena.back().d + c->arena_off;
    size_t at = 0, last = 0;
    for (int i = 0; i < count; ++i) {
        if (bytes[i]) std::memcpy(h + at, srcs[i], bytes[i]);
        outs[i] = d + at;
        last = at + bytes[i];
        at += (bytes[i] + 255) & ~size_t(255);
    }
    // one copy for all segments (the alignment gaps travel along)
    if (last) QH_HIP(hipMemcpyAsync(d, h, last, hipMemcpyHostToDevice, c->stream));
    c->arena_off += need;
}

void* arena_upload(qh_context* c, const void* src, size_t bytes) {
    const size_t need = (bytes + 255) & ~size_t(255);
    if (c->arena.empty() || c->arena_off + need > c->arena.back().cap) {
        qh_context::ArenaChunk ch;
        ch.cap = std::max(need, c->arena.empty() ? c->arena_first : c->arena.back().cap * 2);
        QH_HIP(hipHostMalloc((void**)&ch.h, ch.cap, hipHostMallocDefault));
        if (hipMalloc((void**)&ch.d, ch.cap) != hipSuccess) {
            (void)hipHostFree(ch.h);
            throw Error(QH_ERR_OOM, "descriptor arena allocation failed");
        }
        c->arena.push_back(ch);
        c->arena_off = 0;
    }
    char* h = c->arena.back().h + c->arena_off;
    char* d = c->arena.back().d + c->arena_off;
    if (bytes) {
        std::memcpy(h, src, bytes);
        QH_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream));
    }
    c->arena_off += need;
    return d;
}

void arena_recycle(qh_context* c) {
    if (c->arena.empty() || (c->arena.size() == 1 && c->arena_off == 0)) return;
    // every upload so far has been consumed once the stream is idle
    QH_HIP(hipStreamSynchronize(c->stream));
    if (c->arena.size() > 1) {
        // the last chunk is the largest (capacities double); keep it
        for (size_t i = 0; i + 1 < c->arena.size(); ++i) {
            (void)hipHostFree(c->arena[i].h);
            (void)hipFree(c->arena[i].d);
        }
        c->arena.erase(c->arena.begin(), c->arena.end() - 1);
    }
    c->arena_off = 0;
}

void arena_free(qh_context* c) {
    for (auto& ch : c->arena) {
        (void)hipHostFree(ch.h);
        (void)hipFree(ch.d);
    }
    c->arena.clear();
    c->arena_off = 0;
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
void run_theta(qh_context* c, std::vector<ThetaItem>& items, int precision) {
    if (items.empty()) return;
    const int d = items[0].d;
    const bool mfma = precision != QH_PREC_FP64;
    const int te = mfma ? THETA_MFMA_TILE : theta_tile_extent(d);
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
    if (mfma) launch_theta_mfma(di, dt, (int)tiles.size(), d, precision, c->stream);
    else launch_theta(di, dt, (int)tiles.size(), d, c->stream);
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

// Split-K for an FP64 product whose output tiles alone leave most CUs idle (the
// environment steps of a sweep: a χ × χ output from a contraction over dχ): the K range
// is cut into slices computed as separate partial products into a workspace, then
// added in slice order.  The number of slices depends on the product's own shape only,
// so a product gets the same rounding alone and inside any batch.
static int gemm_slices(const GemmItem& g) {
    const int tiles = ((g.m + GEMM_TM - 1) / GEMM_TM) * ((g.n + GEMM_TN - 1) / GEMM_TN);
    if (tiles >= 64 || g.k < 256) return 1;
    return std::max(1, std::min({8, 256 / tiles, g.k / 64}));
}

void run_gemm(qh_context* c, std::vector<GemmItem>& items, int precision) {
    if (items.empty()) return;
    if (precision == QH_PREC_FP64) {
        // slice length per product: whole K tiles (a slice's K scaling index is offset)
        std::vector<int> ks(items.size(), 0), slices(items.size(), 1);
        size_t ws_elems = 0;
        for (size_t i = 0; i < items.size(); ++i) {
            const GemmItem& g = items[i];
            const int s0 = gemm_slices(g);
            if (s0 <= 1) continue;
            ks[i] = ((g.k + s0 - 1) / s0 + GEMM_TK - 1) / GEMM_TK * GEMM_TK;
            slices[i] = (g.k + ks[i] - 1) / ks[i];
            if (slices[i] > 1) ws_elems += (size_t)slices[i] * g.m * g.n;
        }
        if (ws_elems > 0) {
            cplx* ws = dalloc<cplx>(c, ws_elems);
            std::vector<GemmItem> parts;
            std::vector<GemmReduceItem> red;
            size_t at = 0;
            for (size_t i = 0; i < items.size(); ++i) {
                const GemmItem& g = items[i];
                if (slices[i] <= 1) {
                    parts.push_back(g);
                    continue;
                }
                const size_t mn = (size_t)g.m * g.n;
                red.push_back(GemmReduceItem{ws + at, slices[i], g.c, g.m, g.n, g.ldc, g.alpha, g.beta, g.rowscale,
                                             g.colscale});
                for (int s = 0; s < slices[i]; ++s) {
                    const int k0 = s * ks[i];
                    GemmItem p = g;
                    p.k = std::min(ks[i], g.k - k0);
                    p.dk_off = g.dk_off + k0;
                    p.a = (g.opa == GEMM_N || g.opa == GEMM_CONJ) ? g.a + (size_t)g.lda * k0 : g.a + k0;
                    p.b = (g.opb == GEMM_N || g.opb == GEMM_CONJ) ? g.b + k0 : g.b + (size_t)g.ldb * k0;
                    p.c = ws + at + s * mn;
                    p.ldc = g.m;
                    p.alpha = make_double2(1.0, 0.0);
                    p.beta = make_double2(0.0, 0.0);
                    p.rowscale = nullptr;
                    p.colscale = nullptr;
                    parts.push_back(p);
                }
                at += (size_t)slices[i] * mn;
            }
            run_gemm_tiles(c, parts);
            std::vector<int2> rt;
            for (size_t i = 0; i < red.size(); ++i) {
                const size_t mn = (size_t)red[i].m * red[i].n;
                for (size_t blk = 0; blk * 256 < mn; ++blk) rt.push_back(make_int2((int)i, (int)blk));
            }
            GemmReduceItem* dr;
            int2* drt;
            upload2(c, red, rt, dr, drt);
            launch_gemm_reduce(dr, drt, (int)rt.size(), c->stream);
            dfree(c, ws);
            return;
        }
    }
    run_gemm_tiles(c, items, precision);
}

void run_gemm_tiles(qh_context* c, std::vector<GemmItem>& items, int precision) {
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
    GemmItem* di;
    int2* dt;
    upload2(c, items, tiles, di, dt);
    Timed tm(c, "gemm", flops, bytes);
    if (precision == QH_PREC_FP64) launch_gemm(di, dt, (int)tiles.size(), c->stream);
    else launch_gemm_mfma(di, dt, (int)tiles.size(), precision, c->stream);
}

// ---------------------------------------------------------------------------
// Jacobi SVD
// ---------------------------------------------------------------------------
void run_jacobi(qh_context* c, std::vector<JacItem>& items, double* fro2_out, int block) {
    if (items.empty()) return;
    if (block <= 0) block = c->jac_block > 0 ? c->jac_block : 16;
    const int bc = 2 * block;
    std::vector<int2> id_tiles, tiles;
    int rmax = 1;
    // per item: rotations in the window, converged, tail state, JacItem::own
    constexpr int CS = 3 + JAC_OWN_INTS;
    // then `remaining` (4) and per item the window's rotations and tail state (2)
    const size_t ncount = CS * items.size() + 4 + 2 * items.size();
    int* counters = dalloc<int>(c, ncount);
    QH_HIP(hipMemsetAsync(counters, 0, ncount * sizeof(int), c->stream));
    double* fro2 = fro2_out ? fro2_out : dalloc<double>(c, items.size());
    auto free_fro2 = [&] {
        if (!fro2_out) dfree(c, fro2);
    };
    std::vector<int2> frob_tiles;
    std::vector<int> frob_first(1, 0);
    for (size_t i = 0; i < items.size(); ++i) {
        JacItem& it = items[i];
        int nb = (it.q + block - 1) / block;
        if (nb < 2) nb = 2;
        if (nb & 1) ++nb;
        it.nb = nb;
        it.rot = counters + CS * i;
        it.conv = counters + CS * i + 1;
        it.tail = counters + CS * i + 2;
        it.own = counters + CS * i + 3;
        it.fro2 = fro2 + i;
        const size_t tot = (size_t)it.p * it.q;
        for (size_t k = 0; k * FROB_CHUNK < tot; ++k) frob_tiles.push_back(make_int2((int)i, (int)k));
        frob_first.push_back((int)frob_tiles.size());
        rmax = std::max(rmax, nb - 1);
        for (int k = 0; k < it.q; ++k) id_tiles.push_back(make_int2((int)i, k));
        for (int k = 0; k < nb / 2; ++k) tiles.push_back(make_int2((int)i, k));
    }
    int* remaining = counters + CS * items.size();
    // Relative orthogonality target |x_i^H x_j| <= tol ||x_i|| ||x_j||.  The Gram
    // entries carry O(sqrt(p)·ε) rounding noise, so the target sits a factor 16
    // above it; otherwise noise-triggered rotations never stop.  Columns below
    // 4 sqrt(p) ε ||X||_F are frozen.  Both depend only on the matrix itself, so a
    // matrix gets the same result alone and inside any batch.
    for (JacItem& it : items) {
        const double sp = std::sqrt((double)std::max(it.p, 4));
        it.tol = 16.0 * sp * DBL_EPSILON;
        const double fl = 4.0 * sp * DBL_EPSILON;
        it.floor_rel2 = fl * fl;
    }
    JacItem* di = upload(c, items);
    int2* dti = upload(c, id_tiles);
    int2* dt = upload(c, tiles);
    int2* dtf = upload(c, frob_tiles);
    launch_set_identity_items(di, dti, (int)id_tiles.size(), c->stream);
    double* frob_part = dalloc<double>(c, std::max<size_t>(frob_tiles.size(), 1));
    launch_frob2(di, (int)items.size(), dtf, (int)frob_tiles.size(), frob_part, upload(c, frob_first), c->stream);
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
    std::vector<int> hrot(CS * items.size());
    bool done = false;
    // The tail.  Once one of its own sweeps (after the sorting ones) rotated at most
    // c->jac_tail_switch per mille (700 by default) of its block pairs -- near the quadratic
    // phase, where the next sweep rotates only a few -- an item leaves the rounds (the round
    // kernel decides, per item, so the decision does not depend on the batch), and after the
    // window its remaining sweeps become passes of: one launch testing every (round, pair) of
    // a sweep without rotating; its flagged pairs grouped on the host into sets of disjoint
    // pairs (greedily, in round order); one launch per set, where each pair is tested again on
    // its fresh Gram before it rotates.  A pass whose test flags nothing is a full sweep
    // without a rotation: the item has converged, as in the plain loop.  An item whose first
    // test flags more than TAIL_MAX of its pairs returns to plain sweeps where it left them
    // (its round offset makes its next round the first of the sweep it had not run; it may
    // enter the tail again one sweep later), and so does an item still flagged after
    // TAIL_PASSES passes (for good).
    constexpr int TAIL_MAX = 250;  // per mille of an item's block pairs
    constexpr int TAIL_PASSES = 8;
    static const int kdbg = std::getenv("QH_DEBUG_JACOBI_KERNEL") ? 1 : 0;
    const int tail_switch = (c->jac_tail && !kdbg) ? c->jac_tail_switch : 0;
    int* tail_flags = nullptr;
    std::vector<int> tail_first(items.size());
    size_t tail_total = 0;
    // the test: by 64-column tiles of X^H X at block widths 16 and 32 (flags per block pair
    // I, J of an nb × nb table), per (round, pair) of a sweep at width 8
    const bool tile_check = block != 8;
    for (size_t i = 0; i < items.size(); ++i) {
        tail_first[i] = (int)tail_total;
        const size_t nb = items[i].nb;
        tail_total += tile_check ? nb * nb : (nb - 1) * (nb / 2);
    }
    // the items waiting for the tail (tail = 1) run their passes; next_round is the batch's
    // first round of the next window
    auto run_tail = [&](int next_round) {
        if (!tail_flags) tail_flags = dalloc<int>(c, tail_total);
        const int* d_first = upload(c, tail_first);
        std::vector<int> hcnt(CS * items.size()), hflags(tail_total);
        std::vector<char> had_sets(items.size(), 0);
        auto publish = [&] {  // converged items, tail states, offsets, cleared rotation counts
            // through the descriptor arena (pinned staging that no later upload of this
            // call overwrites), so the host need not wait for the copy
            const int* staged = upload(c, hcnt);
            QH_HIP(hipMemcpyAsync(counters, staged, hcnt.size() * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
        };
        // back to plain sweeps at the sweep it had not run; `again`: the first own sweep at
        // which it may enter the tail once more
        auto leave = [&](size_t i, int again) {
            int* o = &hcnt[CS * i + 3];
            hcnt[CS * i + 2] = 0;
            o[JAC_ROUND_OFF] = next_round - o[JAC_TAIL_AT] * (items[i].nb - 1);
            o[JAC_TAIL_MIN] = again;
        };
        for (int pass = 0;; ++pass) {
            readback(c, hcnt.data(), counters, hcnt.size() * sizeof(int));
            std::vector<int2> ctiles;
            std::vector<int4> ttiles;
            for (size_t i = 0; i < items.size(); ++i) {
                if (hcnt[CS * i + 1] != 0 || hcnt[CS * i + 2] != 1) continue;
                if (c->timing) {  // the rotations of the previous pass's sets
                    c->timers["jacobi_round"].flops += hcnt[CS * i] * upd_f[i];
                    c->timers["jacobi_round"].bytes += hcnt[CS * i] * upd_b[i];
                }
                if (had_sets[i] && hcnt[CS * i] == 0) {
                    // every flagged pair passed the rounds' own test on its fresh Gram and
                    // nothing moved: the two tests differ only in rounding at the margin
                    hcnt[CS * i + 1] = 1;
                    continue;
                }
                const int nb = items[i].nb;
                if (tile_check) {
                    const int nt = (nb * (block / 16) + 3) / 4;  // 64-column tiles
                    for (int ti = 0; ti < nt; ++ti)
                        for (int tj = ti; tj < nt; ++tj) ttiles.push_back(make_int4((int)i, ti, tj, 0));
                } else {
                    for (int e = 0; e < (nb - 1) * (nb / 2); ++e) ctiles.push_back(make_int2((int)i, e));
                }
            }
            if (ctiles.empty() && ttiles.empty()) {
                publish();
                return;
            }
            if (pass == TAIL_PASSES) {
                for (size_t i = 0; i < items.size(); ++i) {
                    hcnt[CS * i] = 0;
                    if (hcnt[CS * i + 1] == 0 && hcnt[CS * i + 2] == 1) leave(i, INT_MAX);
                }
                publish();
                return;
            }
            {
                Timed tm(c, "jacobi_check");
                if (tile_check) {
                    QH_HIP(hipMemsetAsync(tail_flags, 0, tail_total * sizeof(int), c->stream));
                    launch_jacobi_check_tiles(di, upload(c, ttiles), (int)ttiles.size(), tail_flags, d_first, block,
                                              c->stream);
                } else {
                    launch_jacobi_check(di, upload(c, ctiles), (int)ctiles.size(), tail_flags, d_first, block,
                                        c->stream);
                }
            }
            readback(c, hflags.data(), tail_flags, tail_total * sizeof(int));
            // per set: the tiles of the flagged pairs; per item and set: the blocks taken
            std::vector<std::vector<int2>> sets;
            for (size_t i = 0; i < items.size(); ++i) {
                hcnt[CS * i] = 0;
                if (hcnt[CS * i + 1] != 0 || hcnt[CS * i + 2] != 1) continue;
                const int nb = items[i].nb, np = nb / 2, cnt = (nb - 1) * np;
                if (c->timing) {
                    c->timers["jacobi_check"].flops += cnt * gram_f[i];
                    c->timers["jacobi_check"].bytes += cnt * gram_b[i];
                }
                // the flagged block pairs, in round order; from the tiled test, a block with a
                // pair inside it failing and in no flagged pair rotates with its neighbour I ^ 1
                std::vector<int2> fl;
                const int* hf = &hflags[tail_first[i]];
                if (tile_check) {
                    std::vector<char> in(nb, 0), want(nb, 0);
                    for (int I = 0; I < nb; ++I) {
                        want[I] = hf[I * nb + I] != 0;
                        for (int J = I + 1; J < nb; ++J)
                            if (hf[I * nb + J]) in[I] = in[J] = 1;
                    }
                    for (int e = 0; e < cnt; ++e) {
                        int a, b;
                        rr_pair(e / np, e % np, nb, a, b);
                        const int lo = std::min(a, b), hi = std::max(a, b);
                        if (hf[lo * nb + hi] || (hi == (lo ^ 1) && ((want[lo] && !in[lo]) || (want[hi] && !in[hi]))))
                            fl.push_back(make_int2(a, b));
                    }
                } else {
                    for (int e = 0; e < cnt; ++e) {
                        if (!hf[e]) continue;
                        int a, b;
                        rr_pair(e / np, e % np, nb, a, b);
                        fl.push_back(make_int2(a, b));
                    }
                }
                const int flagged = (int)fl.size();
                if (flagged == 0) {
                    hcnt[CS * i + 1] = 1;  // a full sweep without a rotation
                    continue;
                }
                if (pass == 0 && 1000LL * flagged > (long long)TAIL_MAX * cnt) {
                    leave(i, hcnt[CS * i + 3 + JAC_TAIL_AT] + 1);
                    continue;
                }
                had_sets[i] = 1;
                std::vector<std::vector<char>> taken;
                for (const int2 pr : fl) {
                    const int a = pr.x, b = pr.y;
                    size_t k = 0;
                    for (;; ++k) {
                        if (k == taken.size()) taken.emplace_back(nb, 0);
                        if (!taken[k][a] && !taken[k][b]) break;
                    }
                    taken[k][a] = taken[k][b] = 1;
                    if (sets.size() <= k) sets.resize(k + 1);
                    sets[k].push_back(make_int2((int)i, a | (b << 16)));
                }
            }
            publish();
            static const bool tdbg = std::getenv("QH_DEBUG_JACOBI") != nullptr;
            if (tdbg) {
                std::fprintf(stderr, "[jacobi] tail pass %d: %zu sets:", pass, sets.size());
                for (const auto& st : sets) std::fprintf(stderr, " %zu", st.size());
                std::fprintf(stderr, "\n");
            }
            if (sets.empty()) return;
            Timed tm(c, "jacobi_round");
            if (c->timing) c->timers["jacobi_round"].launches += (int)sets.size() - 1;  // Timed counts one
            for (const auto& st : sets)
                launch_jacobi_pairs(di, upload(c, st), (int)st.size(), block, c->jac_max_inner, c->stream);
        }
    };
    // Items still rotating at the sweep cap get a second window of the same length
    // before the call fails; the converged ones are skipped by every launch, so the
    // extra sweeps cost only the stragglers' work (and the launches' latency).
    const int hard_cap = 2 * c->jac_max_sweeps;
    // A sweep's rounds as one HIP graph, captured at the first sweep and replayed by
    // every later one: the rounds read their index as an offset from a counter in
    // device memory that k_conv_update advances by the sweep's length, so each replay
    // runs the next sweep's rounds (the same indices as direct launches).
    const bool use_graph = c->use_graphs && !kdbg;
    int* round_base = use_graph ? remaining + 2 : nullptr;
    // Block width 16: for matrices of at least JAC_DEFER_MIN_Q columns each round's
    // V_IJ <- V_IJ Q runs in the next launch, beside the next round's pairs (two parities of
    // Q slots, 16 KB each, and 4-int headers that start at zero); the window ends with a
    // launch of V workgroups alone.  Smaller matrices update V in their pairs' workgroups.
    // Only where all pairs fit on the device at once (three workgroups per CU) and those
    // matrices hold at least half of them: the V workgroups follow all pairs in a launch,
    // and behind more pairs than the device holds they start last and lengthen every
    // launch (measured: the uneven 512-MPS workload, thousands of pairs, -5 %).
    cplx* vq = nullptr;
    int* vhdr = nullptr;
    int* dvmap = nullptr;
    int nv = 0;
    if (block == 16 && c->jac_defer_v && tiles.size() <= 3 * (size_t)c->num_cus) {
        std::vector<int> vmap(tiles.size(), -1), vlist;
        for (size_t t = 0; t < tiles.size(); ++t)
            if (items[tiles[t].x].q >= JAC_DEFER_MIN_Q) {
                vmap[t] = (int)vlist.size();
                vlist.push_back((int)t);
            }
        nv = 2 * vlist.size() >= tiles.size() ? (int)vlist.size() : 0;
        if (nv > 0) {
            vmap.insert(vmap.end(), vlist.begin(), vlist.end());
            dvmap = upload(c, vmap);
            vq = dalloc<cplx>(c, 2 * (size_t)nv * bc * bc);
            vhdr = dalloc<int>(c, 8 * (size_t)nv);
            QH_HIP(hipMemsetAsync(vhdr, 0, 8 * (size_t)nv * sizeof(int), c->stream));
        }
    }
    auto free_vq = [&] {
        if (vq) dfree(c, vq);
        if (vhdr) dfree(c, vhdr);
    };
    struct SweepGraph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        ~SweepGraph() {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
        }
    } sg;
    for (int sweep = 0; sweep < hard_cap; ++sweep) {
        // read the per-item rotation counters after the window to account work
        {
            // one event pair per window of back-to-back rounds (per-launch events would
            // cost more host time than the short rounds themselves)
            Timed tm(c, "jacobi_round");
            if (c->timing) c->timers["jacobi_round"].launches += rmax - 1 + (vq ? 1 : 0);  // Timed counts one
            if (use_graph) {
                if (!sg.exec) {
                    QH_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                    for (int r = 0; r < rmax; ++r)
                        launch_jacobi_round(di, dt, (int)tiles.size(), r, round_base, block, c->jac_sort_sweeps, 0,
                                            c->jac_max_inner, tail_switch, vq, vhdr, dvmap, nv, c->stream);
                    if (vq)
                        launch_jacobi_vonly(di, dt, (int)tiles.size(), rmax, round_base, vq, vhdr, dvmap, nv,
                                            c->stream);
                    const hipError_t e1 = hipGetLastError();
                    const hipError_t e2 = hipStreamEndCapture(c->stream, &sg.graph);
                    QH_HIP(e1);
                    QH_HIP(e2);
                    QH_HIP(hipGraphInstantiate(&sg.exec, sg.graph, nullptr, nullptr, 0));
                }
                QH_HIP(hipGraphLaunch(sg.exec, c->stream));
            } else {
                for (int r = 0; r < rmax; ++r) {
                    const int g = sweep * rmax + r;
                    launch_jacobi_round(di, dt, (int)tiles.size(), g, nullptr, block, c->jac_sort_sweeps, kdbg,
                                        c->jac_max_inner, tail_switch, vq, vhdr, dvmap, nv, c->stream);
                }
                if (vq)
                    launch_jacobi_vonly(di, dt, (int)tiles.size(), (sweep + 1) * rmax, nullptr, vq, vhdr, dvmap, nv,
                                        c->stream);
            }
        }
        static const bool dbg = std::getenv("QH_DEBUG_JACOBI") != nullptr;
        if (dbg) {
            readback(c, hrot.data(), counters, hrot.size() * sizeof(int));
            std::fprintf(stderr, "[jacobi] sweep %d rmax %d rot:", sweep, rmax);
            for (size_t i = 0; i < items.size(); ++i)
                std::fprintf(stderr, " %d/%d/%d(%dx%d,nb%d)", hrot[CS * i], hrot[CS * i + 1], hrot[CS * i + 2],
                             items[i].p, items[i].q, items[i].nb);
            std::fprintf(stderr, "\n");
        }
        launch_conv_update(di, (int)items.size(), remaining, round_base, rmax, sweep * rmax, c->jac_sort_sweeps,
                           tail_switch, c->stream);
        c->timers["jacobi_sweep"].launches += 1;
        int rem[4] = {0, 0, 0, 0};
        if (c->timing) {  // `remaining` and the window's per-item counts in one read-back
            std::vector<int> hs(4 + 2 * items.size());
            readback(c, hs.data(), remaining, hs.size() * sizeof(int));
            std::copy(hs.begin(), hs.begin() + 4, rem);
            auto& t = c->timers["jacobi_round"];
            for (size_t i = 0; i < items.size(); ++i) {
                const int wrot = hs[4 + 2 * i], wtail = hs[4 + 2 * i + 1];
                if (converged[i] || wtail == 1) continue;
                const double checked = (double)(items[i].nb / 2) * rmax;
                t.flops += checked * gram_f[i] + wrot * upd_f[i];
                t.bytes += checked * gram_b[i] + wrot * upd_b[i];
                if (wrot == 0) converged[i] = 1;
            }
        } else {
            readback(c, rem, remaining, sizeof(rem));
        }
        if (rem[1] > 0) {
            std::vector<int> h(CS * items.size());
            readback(c, h.data(), counters, h.size() * sizeof(int));
            std::string msg = "non-finite entries in the matrix to decompose:";
            for (size_t i = 0; i < items.size(); ++i)
                if (h[CS * i + 1] < 0) msg += " item " + std::to_string(i);
            dfree(c, counters);
            if (tail_flags) dfree(c, tail_flags);
            free_fro2();
            dfree(c, frob_part);
            free_vq();
            throw Error(QH_ERR_NUMERIC, msg);
        }
        if (rem[3] > 0) {  // items waiting for the tail passes
            run_tail((sweep + 1) * rmax);
            readback(c, hrot.data(), counters, hrot.size() * sizeof(int));
            bool left = false;  // an item back in plain sweeps, or still rotating in them
            for (size_t i = 0; i < items.size(); ++i) left = left || hrot[CS * i + 1] == 0;
            if (!left) {
                done = true;
                break;
            }
            continue;
        }
        if (rem[0] == 0) {
            done = true;
            break;
        }
    }
    if (!done) {
        std::vector<int> h(CS * items.size());
        readback(c, h.data(), counters, h.size() * sizeof(int));
        std::string msg = "Jacobi SVD did not converge within " + std::to_string(hard_cap) + " sweeps:";
        if (const char* dump = std::getenv("QH_DEBUG_JACOBI_DUMP")) {
            for (size_t i = 0; i < items.size(); ++i)
                if (!h[CS * i + 1]) {
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
            if (!h[CS * i + 1]) {
                msg += " item " + std::to_string(i) + " (" + std::to_string(items[i].p) + "x" +
                       std::to_string(items[i].q) + ")";
                ++shown;
            }
        dfree(c, counters);
        if (tail_flags) dfree(c, tail_flags);
        free_fro2();
        dfree(c, frob_part);
        free_vq();
        throw Error(QH_ERR_NOT_CONVERGED, msg);
    }
    dfree(c, counters);
    if (tail_flags) dfree(c, tail_flags);
    free_fro2();
    dfree(c, frob_part);
    free_vq();
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
    GatherItem* di;
    int2* dt;
    upload2(c, items, tiles, di, dt);
    Timed tm(c, "gather", 0.0, bytes);
    launch_gather_items(di, dt, (int)tiles.size(), c->stream);
}


// ---------------------------------------------------------------------------
// Householder QR (preconditioner)
// ---------------------------------------------------------------------------
void run_qr_factor(qh_context* c, const std::vector<QrJob>& jobs) {
    if (jobs.empty()) return;
    std::vector<QrItem> items(jobs.size());
    std::vector<QrApplyItem> app(jobs.size());
    // trailing tiles of a panel step: tile 0 is the next panel's columns, tiles >= 1
    // the columns beyond it
    std::vector<int2> next_tiles, rest_tiles;
    int nmax = 0, mmax = 0;
    const int tc = qr_tile_cols();
    double flops = 0;
    for (size_t i = 0; i < jobs.size(); ++i) {
        const QrJob& j = jobs[i];
        mmax = std::max(mmax, j.m);
        items[i] = QrItem{j.a, j.m, j.n, j.lda, j.v, j.t};
        app[i] = QrApplyItem{j.v, j.t, j.m, j.n, j.a, j.lda, j.n, 1};
        nmax = std::max(nmax, j.n);
        const int ntile = (std::max(j.n - QR_NB, 0) + tc - 1) / tc;
        if (ntile > 0) next_tiles.push_back(make_int2((int)i, 0));
        for (int k = 1; k < ntile; ++k) rest_tiles.push_back(make_int2((int)i, k));
        flops += 32.0 / 3.0 * j.n * (double)j.n * (3.0 * j.m - j.n) / 2.0;
    }
    QrItem* di = upload(c, items);
    QrApplyItem* da = upload(c, app);
    int2* dn = upload(c, next_tiles);
    int2* dr = upload(c, rest_tiles);
    Timed tm(c, "qr", flops, 0.0);
    if (mmax > 512) {
        // panels of up to 1024 rows in registers need two rows per thread, which the
        // fused look-ahead step (one row per thread of 512) does not hold: panel, then
        // trailing update
        for (int j0 = 0; j0 < nmax; j0 += QR_NB) {
            launch_qr_panel_tall(di, (int)items.size(), j0, c->stream);
            launch_qr_apply(da, dn, (int)next_tiles.size(), j0, 1, c->stream);
            launch_qr_apply(da, dr, (int)rest_tiles.size(), j0, 1, c->stream);
        }
        return;
    }
    // look-ahead: after panel j0, one launch updates the next panel's columns and
    // factors that panel (a workgroup per item) while the others apply panel j0 to the
    // columns beyond it
    launch_qr_panel(di, (int)items.size(), 0, c->stream);
    for (int j0 = 0; j0 + QR_NB < nmax; j0 += QR_NB)
        launch_qr_step(di, (int)items.size(), da, dr, (int)rest_tiles.size(), j0, c->stream);
}

void run_qr_apply_q(qh_context* c, const std::vector<QrApplyJob>& jobs) {
    if (jobs.empty()) return;
    std::vector<QrApplyItem> app(jobs.size());
    std::vector<int2> tiles;
    int nmax = 0;
    const int tc = qr_tile_cols();
    for (size_t i = 0; i < jobs.size(); ++i) {
        const QrJob& q = *jobs[i].qr;
        app[i] = QrApplyItem{q.v, q.t, q.m, q.n, jobs[i].y, q.m, jobs[i].ncols, 0};
        nmax = std::max(nmax, q.n);
        for (int k = 0; k < (jobs[i].ncols + tc - 1) / tc; ++k) tiles.push_back(make_int2((int)i, k));
    }
    if (tiles.empty() || nmax <= 0) return;
    QrApplyItem* da = upload(c, app);
    int2* dt = upload(c, tiles);
    Timed tm(c, "qr_apply");
    // column tiles are independent: one launch, each workgroup applies every panel
    launch_qr_form(da, dt, (int)tiles.size(), c->stream);
}

// F <- F (3I - F^H F) / 2 for every job (in place; zero columns stay zero)
void polar_refine(qh_context* c, const std::vector<PolarJob>& jobs) {
    if (jobs.empty()) return;
    std::vector<cplx*> G(jobs.size()), F2(jobs.size());
    std::vector<GemmItem> gg, gf;
    for (size_t q = 0; q < jobs.size(); ++q) {
        const PolarJob& j = jobs[q];
        G[q] = dalloc<cplx>(c, (size_t)j.cols * j.cols);
        F2[q] = dalloc<cplx>(c, (size_t)j.rows * j.cols);
        gg.push_back(gemm_item(j.f, GEMM_C, j.rows, j.f, GEMM_N, j.rows, G[q], j.cols, j.cols, j.cols, j.rows));
        GemmItem h = gemm_item(j.f, GEMM_N, j.rows, G[q], GEMM_N, j.cols, F2[q], j.rows, j.rows, j.cols, j.cols);
        h.alpha = make_double2(-0.5, 0.0);
        h.beta = make_double2(1.5, 0.0);
        gf.push_back(h);
    }
    run_gemm(c, gg);
    for (size_t q = 0; q < jobs.size(); ++q)
        QH_HIP(hipMemcpyAsync(F2[q], jobs[q].f, (size_t)jobs[q].rows * jobs[q].cols * sizeof(cplx),
                              hipMemcpyDeviceToDevice, c->stream));
    run_gemm(c, gf);
    for (size_t q = 0; q < jobs.size(); ++q) {
        QH_HIP(hipMemcpyAsync(jobs[q].f, F2[q], (size_t)jobs[q].rows * jobs[q].cols * sizeof(cplx),
                              hipMemcpyDeviceToDevice, c->stream));
        dfree(c, G[q]);
        dfree(c, F2[q]);
    }
}

// QR preconditioning passes for a matrix with q columns: the tuned value of its size
// class (qh_ctx_autotune) if there is one, else the context's setting
// the tuned width of the batch's largest size class, else the context's setting,
// else the automatic choice
int block_for(const qh_context* c, int qmax) {
    if (!c->tuned_block.empty()) {
        auto it = c->tuned_block.lower_bound(qmax);
        if (it == c->tuned_block.end()) --it;
        return it->second;
    }
    if (c->jac_block > 0) return c->jac_block;
    return qmax >= 2048 ? 32 : 16;
}

int qr_passes_for(const qh_context* c, int q) {
    int passes = c->jac_precondition;
    if (!c->tuned_passes.empty()) {
        auto it = c->tuned_passes.lower_bound(q);  // the smallest class >= q
        if (it == c->tuned_passes.end()) --it;
        passes = it->second;
    }
    return std::min(std::max(passes, 0), QR_MAX_PASSES);
}

int SvdTask::kmax() const {
    const int k = std::min(mr, mc);
    if (truncate && maxdim >= 0) return std::min(k, maxdim);
    return k;
}

void run_svd_tasks(qh_context* c, std::vector<SvdTask>& tasks) {
    if (tasks.empty()) return;
    const size_t nt = tasks.size();
    int qmax = 0;
    for (const SvdTask& t : tasks) qmax = std::max(qmax, std::min(t.mr, t.mc));
    const int bw = block_for(c, qmax);
    // workspace layout per task
    size_t bytes = 0;
    auto take = [&](size_t b) {
        size_t o = bytes;
        bytes += (b + 255) & ~size_t(255);
        return o;
    };
    struct Off {
        size_t v, s, p, l, ux, vx, tmp;
        size_t qv[QR_MAX_PASSES + 1], qt[QR_MAX_PASSES + 1], y[QR_MAX_PASSES + 1];
        bool pre;
        int np;  // QR preconditioning passes of this task (0 without preconditioning)
        int p_, q_, k_;
    };
    std::vector<Off> off(nt);
    for (size_t i = 0; i < nt; ++i) {
        const SvdTask& t = tasks[i];
        Off& o = off[i];
        o.p_ = std::max(t.mr, t.mc);
        o.q_ = std::min(t.mr, t.mc);
        o.k_ = t.kmax();
        const size_t p = o.p_, q = o.q_;
        // QR preconditioning above one block pair of the narrower widths (q > 16 or
        // 32); the Jacobi's relative accuracy on row-graded θ needs it however wide
        // the blocks are
        o.np = (int)q > std::min(2 * bw, 32) ? qr_passes_for(c, (int)q) : 0;
        o.pre = o.np > 0;
        const int npass = o.np;
        o.v = take(q * q * sizeof(cplx));
        o.s = take(q * sizeof(double));
        o.p = take(q * sizeof(int));
        o.l = take(q * sizeof(double));
        if (o.pre) {
            const size_t tq = ((q + QR_NB - 1) / QR_NB) * QR_NB * QR_NB;
            for (int j = 1; j <= npass; ++j) {
                o.qv[j] = take((j == 1 ? p : q) * q * sizeof(cplx));
                o.qt[j] = take(tq * sizeof(cplx));
                o.y[j] = take(q * q * sizeof(cplx));
            }
            o.vx = take(q * std::max(o.k_, 1) * sizeof(cplx));
        }
    }
    // The buffers that start as zeros (ux and tmp) sit together, cleared by one fill.
    // (This placement of the per-task buffers is the measured one: with vx moved into
    // the zero-filled run, the same Jacobi rounds took 9 % longer on MI355X -- the X and
    // V buffers of the batch's matrices then sit at other strides in HBM.)
    const size_t zero_begin = bytes;
    for (size_t i = 0; i < nt; ++i) {
        Off& o = off[i];
        if (!o.pre) continue;
        const size_t p = o.p_, q = o.q_, k = std::max(o.k_, 1);
        o.ux = take(p * k * sizeof(cplx));
        if (o.np >= 3) o.tmp = take(q * k * sizeof(cplx));
    }
    const size_t zero_end = bytes;
    const size_t off_k = take(nt * sizeof(int));
    const size_t off_d = take(nt * sizeof(double));
    const size_t off_f = take(nt * sizeof(double));
    char* ws = static_cast<char*>(dalloc_bytes(c, bytes));
    auto P = [&](size_t o) { return reinterpret_cast<cplx*>(ws + o); };
    int* d_kept = reinterpret_cast<int*>(ws + off_k);
    double* d_disc = reinterpret_cast<double*>(ws + off_d);

    // ---- column pivoting of the first pass: X P = Q1 R1 with the columns of X in order
    // of decreasing norm (R1 then has a graded diagonal, and the Jacobi needs fewer
    // sweeps).  The passes and the Jacobi then factor X P, so X's right factor is P
    // times theirs: its output rows are gathered through P's inverse.
    std::vector<cplx*> piv_copy(nt, nullptr);
    std::vector<const int*> piv_inv(nt, nullptr);
    int* piv_buf = nullptr;
    if (c->pivot) {
        size_t ints = 0;
        for (const Off& o : off)
            if (o.pre && o.q_ <= PIVOT_MAXQ) ints += 2 * (size_t)o.q_ + 1;
        if (ints > 0) {
            piv_buf = dalloc<int>(c, ints);
            std::vector<PivotItem> pv;
            std::vector<GatherItem> pg;
            size_t at = 0;
            for (size_t i = 0; i < nt; ++i) {
                const Off& o = off[i];
                if (!o.pre || o.q_ > PIVOT_MAXQ) continue;
                int* perm = piv_buf + at;
                int* inv = perm + o.q_;
                int* count = inv + o.q_;
                at += 2 * (size_t)o.q_ + 1;
                piv_copy[i] = dalloc<cplx>(c, (size_t)o.p_ * o.q_);
                piv_inv[i] = inv;
                pv.push_back(PivotItem{tasks[i].x, o.p_, o.q_, perm, inv, count});
                GatherItem g{};
                g.src = tasks[i].x; g.ld_src = o.p_; g.rows = o.p_; g.perm = perm; g.kept = count;
                g.rowscale_mod = 1; g.rowscale_div = 1; g.out = piv_copy[i]; g.ld_out = o.p_; g.max_cols = o.q_;
                pg.push_back(g);
            }
            launch_pivot_order(upload(c, pv), (int)pv.size(), c->stream);
            run_gather(c, pg);
        }
    }

    // ---- QR preconditioning: X = Q1 R1 and R_{j-1}^H = Q_j R_j for j = 2..npass;
    // the Jacobi runs on Y = R_npass^H (lower triangular), y[j] = R_j^H
    std::vector<std::array<QrJob, QR_MAX_PASSES + 1>> qr(nt);
    int max_np = 0;
    for (const Off& o : off) max_np = std::max(max_np, o.np);
    for (int j = 1; j <= max_np; ++j) {
        std::vector<QrJob> f;
        for (size_t i = 0; i < nt; ++i) {
            const Off& o = off[i];
            if (o.np < j) continue;
            cplx* x1 = piv_copy[i] ? piv_copy[i] : tasks[i].x;
            qr[i][j] = (j == 1) ? QrJob{x1, o.p_, o.q_, o.p_, P(o.qv[1]), P(o.qt[1])}
                                : QrJob{P(o.y[j - 1]), o.q_, o.q_, o.q_, P(o.qv[j]), P(o.qt[j])};
            f.push_back(qr[i][j]);
        }
        if (f.empty()) break;
        run_qr_factor(c, f);
        std::vector<UpperHItem> uh;
        std::vector<int2> ut;
        for (size_t i = 0; i < nt; ++i)
            if (off[i].np >= j) {
                const int q = off[i].q_;
                for (size_t b = 0; b * 256 < (size_t)q * q; ++b) ut.push_back(make_int2((int)uh.size(), (int)b));
                uh.push_back(UpperHItem{qr[i][j].a, qr[i][j].m, q, P(off[i].y[j])});
            }
        launch_upper_h_batch(upload(c, uh), upload(c, ut), (int)ut.size(), c->stream);
    }
    for (cplx* t : piv_copy)
        if (t) dfree(c, t);

    std::vector<JacItem> jac(nt);
    std::vector<SelectItem> sel(nt);
    for (size_t i = 0; i < nt; ++i) {
        const SvdTask& t = tasks[i];
        const Off& o = off[i];
        jac[i] = JacItem{};
        jac[i].x = o.pre ? P(o.y[o.np]) : t.x;
        jac[i].v = P(o.v);
        jac[i].p = o.pre ? o.q_ : o.p_;
        jac[i].q = o.q_;
        SelectItem& s = sel[i];
        s = SelectItem{};
        s.x = jac[i].x;
        s.p = jac[i].p;
        s.q = o.q_;
        s.sigma_sorted = reinterpret_cast<double*>(ws + o.s);
        s.perm = reinterpret_cast<int*>(ws + o.p);
        s.lam_out = t.lam_dst ? t.lam_dst : reinterpret_cast<double*>(ws + o.l);
        s.kept = d_kept + i;
        s.discarded = d_disc + i;
        s.threshold = t.threshold;
        s.maxdim = t.maxdim;
        s.truncate = t.truncate;
        s.renormalize = t.renormalize;
    }
    double* d_fro2 = reinterpret_cast<double*>(ws + off_f);
    run_jacobi(c, jac, d_fro2, bw);
    run_select(c, sel);
    // kept columns the Jacobi left frozen (at or below its deflation floor) get an
    // orthonormal completion, as LAPACK's factors are orthonormal however small σ is
    {
        std::vector<CompleteItem> comp(nt);
        std::vector<int2> ct;
        for (size_t i = 0; i < nt; ++i) {
            comp[i] = CompleteItem{jac[i].x, jac[i].p, sel[i].perm, sel[i].sigma_sorted, sel[i].lam_out,
                                   sel[i].kept, d_fro2 + i, jac[i].floor_rel2};
            for (int j = 0; j < off[i].k_; ++j) ct.push_back(make_int2((int)i, j));
        }
        CompleteItem* dcomp = upload(c, comp);
        launch_complete_project(dcomp, upload(c, ct), (int)ct.size(), c->stream);
        int* d_cnt = dalloc<int>(c, 2 * nt);
        launch_complete_count(dcomp, (int)nt, d_cnt, c->stream);
        std::vector<int> cnt(2 * nt);
        readback(c, cnt.data(), d_cnt, cnt.size() * sizeof(int));
        // the projected starts orthonormalised among themselves: Householder QR of the
        // frozen block (gathered to p × f), its Q written back (times σ where σ > 0)
        std::vector<size_t> fi;
        for (size_t i = 0; i < nt; ++i)
            if (cnt[2 * i + 1] > cnt[2 * i]) fi.push_back(i);
        if (!fi.empty()) {
            launch_complete_zero_tiny(dcomp, d_cnt, (int)nt, c->stream);
            std::vector<int> fcount(fi.size());
            for (size_t k = 0; k < fi.size(); ++k) fcount[k] = cnt[2 * fi[k] + 1] - cnt[2 * fi[k]];
            const int* d_f = upload(c, fcount);
            std::vector<QrJob> jobs(fi.size());
            std::vector<cplx*> bufs;
            std::vector<GatherItem> gz;
            std::vector<ScatterItem> sq;
            std::vector<int2> st;
            for (size_t k = 0; k < fi.size(); ++k) {
                const size_t i = fi[k];
                const int p = jac[i].p, f = fcount[k], first = cnt[2 * i];
                cplx* z = dalloc<cplx>(c, (size_t)p * f);
                cplx* v = dalloc<cplx>(c, (size_t)p * f);
                cplx* t = dalloc<cplx>(c, (size_t)((f + QR_NB - 1) / QR_NB) * QR_NB * QR_NB);
                cplx* q = dalloc<cplx>(c, (size_t)p * f);
                bufs.insert(bufs.end(), {z, v, t, q});
                GatherItem g{};
                g.src = jac[i].x; g.ld_src = p; g.rows = p; g.perm = sel[i].perm + first; g.kept = d_f + k;
                g.rowscale_mod = 1; g.rowscale_div = 1; g.out = z; g.ld_out = p; g.max_cols = f;
                gz.push_back(g);
                jobs[k] = QrJob{z, p, f, p, v, t};
                launch_eye_rect(q, p, f, c->stream);
                sq.push_back(ScatterItem{jac[i].x, p, sel[i].perm + first, sel[i].sigma_sorted + first, d_f + k, q,
                                         p, 1});
                for (int j = 0; j < f; ++j) st.push_back(make_int2((int)k, j));
            }
            run_gather(c, gz);
            run_qr_factor(c, jobs);
            std::vector<QrApplyJob> form;
            for (size_t k = 0; k < fi.size(); ++k) form.push_back(QrApplyJob{&jobs[k], bufs[4 * k + 3], fcount[k]});
            run_qr_apply_q(c, form);
            launch_scatter_cols(upload(c, sq), upload(c, st), (int)st.size(), c->stream);
            for (cplx* b : bufs) dfree(c, b);
        }
        dfree(c, d_cnt);
    }
    // polar refinement where asked (see SvdTask::refine).  Without preconditioning the
    // factors are the kept columns of X (divided by σ) and of V, refined in dense
    // copies and written back; with it, ux and vx below, after the Q factors.
    {
        std::vector<PolarJob> jobs;
        std::vector<cplx*> tmp;
        std::vector<GatherItem> gi;
        std::vector<ScatterItem> si;
        std::vector<int2> st;
        for (size_t i = 0; i < nt; ++i) {
            if (!tasks[i].refine || off[i].pre) continue;
            const int k = std::max(off[i].k_, 1);
            for (int side = 0; side < 2; ++side) {
                const int rows = side == 0 ? jac[i].p : jac[i].q;
                cplx* f = dalloc<cplx>(c, (size_t)rows * k);
                tmp.push_back(f);
                launch_fill(reinterpret_cast<double*>(f), 2 * (size_t)rows * k, 0.0, c->stream);
                GatherItem g{};
                g.src = side == 0 ? jac[i].x : jac[i].v; g.ld_src = rows; g.rows = rows; g.perm = sel[i].perm;
                g.sigma_sorted = sel[i].sigma_sorted; g.divide_sigma = side == 0 ? 1 : 0; g.kept = sel[i].kept;
                g.rowscale_mod = 1; g.rowscale_div = 1; g.out = f; g.ld_out = rows; g.max_cols = k;
                gi.push_back(g);
                jobs.push_back(PolarJob{f, rows, k});
                for (int j = 0; j < k; ++j) st.push_back(make_int2((int)si.size(), j));
                si.push_back(ScatterItem{side == 0 ? jac[i].x : jac[i].v, rows, sel[i].perm, sel[i].sigma_sorted,
                                         sel[i].kept, f, rows, side == 0 ? 1 : 0});
            }
        }
        if (!jobs.empty()) {
            run_gather(c, gi);
            polar_refine(c, jobs);
            launch_scatter_cols(upload(c, si), upload(c, st), (int)st.size(), c->stream);
            for (cplx* f : tmp) dfree(c, f);
        }
    }
    // per task: the factors to apply on each side, innermost first
    std::vector<std::vector<int>> uq(nt), vq(nt);
    size_t stages = 0;
    for (size_t i = 0; i < nt; ++i) {
        for (int j = off[i].np; j >= 1; --j) (j % 2 ? uq[i] : vq[i]).push_back(j);
        stages = std::max(stages, std::max(uq[i].size(), vq[i].size()));
    }
    if (zero_end > zero_begin)
        launch_fill(reinterpret_cast<double*>(ws + zero_begin), (zero_end - zero_begin) / sizeof(double), 0.0,
                    c->stream);
    // the polar refinement reads every column of vx: the ones beyond the kept must be zero
    for (size_t i = 0; i < nt; ++i)
        if (off[i].pre && tasks[i].refine)
            launch_fill(reinterpret_cast<double*>(P(off[i].vx)), 2 * (size_t)off[i].q_ * std::max(off[i].k_, 1), 0.0,
                        c->stream);
    std::vector<GatherItem> pre_g;
    std::vector<std::vector<GatherItem>> to_ux(stages + 1);  // by the stage that first needs ux
    for (size_t i = 0; i < nt; ++i) {
        const Off& o = off[i];
        if (!o.pre) continue;
        const bool even = o.np % 2 == 0;
        GatherItem g{};
        g.src = even ? jac[i].x : jac[i].v; g.ld_src = o.q_; g.rows = o.q_; g.perm = sel[i].perm;
        g.sigma_sorted = sel[i].sigma_sorted;
        g.divide_sigma = even ? 1 : 0; g.kept = sel[i].kept; g.rowscale_mod = 1; g.rowscale_div = 1;
        g.conj_transpose = 0; g.max_cols = o.k_;
        if (uq[i].size() > 1) {  // the q × q factors act on a q-row copy first
            g.out = P(o.tmp); g.ld_out = o.q_;
            GatherItem m = g;
            m.src = P(o.tmp); m.perm = nullptr; m.divide_sigma = 0; m.out = P(o.ux); m.ld_out = o.p_;
            to_ux[uq[i].size() - 1].push_back(m);  // Q1 is the last U-side factor
        } else {
            g.out = P(o.ux); g.ld_out = o.p_;
        }
        pre_g.push_back(g);
        GatherItem h{};
        h.src = even ? jac[i].v : jac[i].x; h.ld_src = o.q_; h.rows = o.q_; h.perm = sel[i].perm;
        h.sigma_sorted = sel[i].sigma_sorted;
        h.divide_sigma = even ? 0 : 1; h.kept = sel[i].kept; h.rowscale_mod = 1; h.rowscale_div = 1;
        h.out = P(o.vx); h.conj_transpose = 0; h.max_cols = o.k_;
        pre_g.push_back(h);
    }
    run_gather(c, pre_g);
    for (size_t r = 0; r < stages; ++r) {
        run_gather(c, to_ux[r]);
        std::vector<QrApplyJob> app;
        for (size_t i = 0; i < nt; ++i) {
            const Off& o = off[i];
            if (!o.pre) continue;
            if (r < uq[i].size()) app.push_back(QrApplyJob{&qr[i][uq[i][r]], P(uq[i][r] == 1 ? o.ux : o.tmp), o.k_});
            if (r < vq[i].size()) app.push_back(QrApplyJob{&qr[i][vq[i][r]], P(o.vx), o.k_});
        }
        run_qr_apply_q(c, app);
    }
    {
        std::vector<PolarJob> jobs;
        for (size_t i = 0; i < nt; ++i)
            if (tasks[i].refine && off[i].pre) {
                jobs.push_back(PolarJob{P(off[i].ux), off[i].p_, std::max(off[i].k_, 1)});
                jobs.push_back(PolarJob{P(off[i].vx), off[i].q_, std::max(off[i].k_, 1)});
            }
        polar_refine(c, jobs);
    }

    // ---- outputs
    std::vector<GatherItem> gat;
    for (size_t i = 0; i < nt; ++i) {
        const SvdTask& t = tasks[i];
        const Off& o = off[i];
        const bool xh = t.mr < t.mc;  // X holds M^H
        for (const SvdOut& so : t.outs) {
            GatherItem g{};
            const bool from_x = (so.factor == 0) != xh;  // the factor is X's left factor
            g.rows = (so.factor == 0) ? t.mr : t.mc;
            g.kept = sel[i].kept;
            g.sigma_sorted = sel[i].sigma_sorted;
            g.rowscale = so.rowscale;
            g.rowscale_mod = so.rowscale_mod > 0 ? so.rowscale_mod : 1;
            g.rowscale_div = so.rowscale_div > 0 ? so.rowscale_div : 1;
            g.out = so.dst;
            g.conj_transpose = so.conj_transpose;
            g.max_cols = o.k_;
            if (o.pre) {
                g.src = from_x ? P(o.ux) : P(o.vx);
                g.ld_src = from_x ? o.p_ : o.q_;
                g.perm = nullptr;
                g.divide_sigma = 0;
                if (!from_x) g.row_src = piv_inv[i];
            } else {
                g.src = from_x ? t.x : jac[i].v;
                g.ld_src = from_x ? o.p_ : o.q_;
                g.perm = sel[i].perm;
                g.divide_sigma = from_x ? 1 : 0;
            }
            gat.push_back(g);
        }
    }
    run_gather(c, gat);
    std::vector<int> kept(nt);
    std::vector<double> disc(nt);
    readback(c, kept.data(), d_kept, kept.size() * sizeof(int));
    readback(c, disc.data(), d_disc, disc.size() * sizeof(double));
    for (size_t i = 0; i < nt; ++i) {
        tasks[i].kept = kept[i];
        tasks[i].discarded = disc[i];
    }
    if (piv_buf) dfree(c, piv_buf);
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
    keep(m->ctx, buf);
    m->site[site - 1] = buf;
}

void replace_lam(qh_mps* m, int bond, double* buf) {
    dfree(m->ctx, m->lam[bond - 1]);
    keep(m->ctx, buf);
    m->lam[bond - 1] = buf;
}

}  // namespace qh
