// MPO × MPS on the device: upload of an open MPO (`MPO(arrays)`,
// src/Ansatz/Chain.jl:20,133-172), exact application (bond χ·D) and variational
// compression of H|ψ⟩ into an MPS of given bond dimensions.  The reference has
// no apply/compress routine (it would `merge` the two networks, src/Quantum.jl:
// 330-348, and contract); the algorithm here is the standard one-site fitting
// sweep, restated by the oracle (oracle/mpo.py) for parity.
//
// Index names: ψ bonds b, c (left, right), φ bonds a, x, MPO bonds w, v,
// physical out o, in i.  Device layouts (first index fastest):
//   ψ site A[b, i, c], φ site B[a, o, x], MPO W[w, o, i, v],
//   environments L[b, a, w] (left of a site) and R[c, x, v] (right of it).
// Every contraction is a batch of plain column-major GEMMs on sub-matrices:
//   Y[b, x, i, v] = Σ_c A[b, i, c] R[c, x, v]                (d·D GEMMs)
//   Z[(b,x), (o,w)] = Y[(b,x), (i,v)] · Wr[(i,v), (o,w)]     (1 GEMM, K = d·D)
//   R'[b, a, w] = Σ_o Z[b, x, o, w] · B[a, o, x]^H           (D GEMMs per o, accumulated)
//   M[a, o, x] = Σ_w L[b, a, w]^T Z[b, x, o, w]               (d GEMMs per w, accumulated)
// and the mirror images for L.  `precision` selects FP64 (vector ALUs) or
// FP32 / bf16x3 (MFMA) for these GEMMs; orthonormalisation is FP64 Householder QR.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "qh_runtime.h"

struct qh_mpo {
    qh_context* ctx = nullptr;
    int n = 0, d = 0;
    bool periodic = false;            // `pMPO`: D[0] = D[n] = the wrap bond
    std::vector<int> D;               // n+1 entries, D[0] = D[n] = 1 when open
    std::vector<qh::cplx*> w;         // W[w, o, i, v]
    std::vector<qh::cplx*> wr;        // (d·D_right) × (d·D_left): Wr[(i + d v), (o + d w)]
    std::vector<qh::cplx*> wl;        // (d·D_left) × (d·D_right): Wl[(i + d w), (o + d v)]
};

namespace qh {
void launch_mpo_apply(const cplx* w, const cplx* a, cplx* out, int chb, int chc, int dl, int dr, int d, hipStream_t s);
void launch_eye_rect(cplx* p, int m, int n, hipStream_t s);
void launch_dot(const cplx* x, const cplx* y, size_t n, cplx* out, hipStream_t s);
}

using namespace qh;

namespace {

using zc = cplx;

void check_mpo(qh_context* c, const qh_mpo* h) {
    if (!h) throw Error(QH_ERR_ARGUMENT, "null MPO handle");
    if (h->ctx != c) throw Error(QH_ERR_ARGUMENT, "MPO belongs to another context");
}

GemmItem sub_gemm(const cplx* a, int opa, int lda, const cplx* b, int opb, int ldb, cplx* c, int ldc, int m, int n,
                  int k, double beta) {
    GemmItem g = gemm_item(a, opa, lda, b, opb, ldb, c, ldc, m, n, k);
    g.beta = make_double2(beta, 0.0);
    return g;
}

// Λ vectors of φ go into the site on their right; afterwards φ has none
void absorb_lambdas(qh_context* c, qh_mps* m) {
    for (int b = 1; b <= m->n - 1; ++b) {
        if (!m->lam[b - 1]) continue;
        launch_scale_bond(m->site[b], m->chi[b], m->d, m->chi[b + 1], m->lam[b - 1], 0, 0, 0.0, c->stream);
        replace_lam(m, b, nullptr);
    }
}

// Orthonormal basis of the column space of X (rows × cols, rows >= cols), in place.
void orthonormalize_columns(qh_context* c, cplx* x, int rows, int cols) {
    const size_t tq = (size_t)((cols + QR_NB - 1) / QR_NB) * QR_NB * QR_NB;
    cplx* v = dalloc<cplx>(c, (size_t)rows * cols);
    cplx* t = dalloc<cplx>(c, tq);
    std::vector<QrJob> jobs{QrJob{x, rows, cols, rows, v, t}};
    run_qr_factor(c, jobs);
    launch_eye_rect(x, rows, cols, c->stream);
    std::vector<QrApplyJob> app{QrApplyJob{&jobs[0], x, cols}};
    run_qr_apply_q(c, app);
    dfree(c, v);
    dfree(c, t);
}

// site tensor B (χl × d × χr): left-orthonormal (rows (l,o)) or right-orthonormal
// (rows l, columns (o,r)) basis of its current span
void left_orthonormalize(qh_context* c, cplx* b, int chil, int d, int chir) {
    if (chil * d < chir) throw Error(QH_ERR_DIMENSION, "bond dimension exceeds the left dimension of the site");
    orthonormalize_columns(c, b, chil * d, chir);
}
void right_orthonormalize(qh_context* c, cplx* b, int chil, int d, int chir) {
    const int cols = d * chir;
    if (cols < chil) throw Error(QH_ERR_DIMENSION, "bond dimension exceeds the right dimension of the site");
    cplx* bh = dalloc<cplx>(c, (size_t)chil * cols);
    launch_conj_transpose(b, chil, cols, bh, c->stream);  // (o r) × l
    orthonormalize_columns(c, bh, cols, chil);
    launch_conj_transpose(bh, cols, chil, b, c->stream);
    dfree(c, bh);
}

struct Fit {
    qh_context* c;
    const qh_mpo* h;
    const qh_mps* psi;
    qh_mps* phi;
    int prec;
    int n, d;
    std::vector<cplx*> L, R;  // L[i]: left of site i+1 (i = 0..n-1), R[i]: right of site i (i = 1..n)

    int cb(int bond) const { return psi->chi[bond]; }   // ψ bond dims
    int ca(int bond) const { return phi->chi[bond]; }   // φ bond dims
    int dw(int bond) const { return h->D[bond]; }

    cplx* ones() {
        cplx* p = dalloc<cplx>(c, 1);
        launch_eye_rect(p, 1, 1, c->stream);
        return p;
    }

    // Z[(b,x),(o,w)] for site i from R[i] (right environment of site i)
    cplx* z_from_right(int i) {
        const int b = cb(i - 1), cc = cb(i), x = ca(i), v = dw(i), w = dw(i - 1);
        const cplx* A = psi->site[i - 1];
        const cplx* Ri = R[i];
        cplx* Y = dalloc<cplx>(c, (size_t)b * x * d * v);
        std::vector<GemmItem> g;
        for (int vv = 0; vv < v; ++vv)
            for (int ii = 0; ii < d; ++ii)
                g.push_back(sub_gemm(A + (size_t)b * ii, GEMM_N, b * d, Ri + (size_t)cc * x * vv, GEMM_N, cc,
                                     Y + (size_t)b * x * (ii + d * vv), b, b, x, cc, 0.0));
        run_gemm(c, g, prec);
        cplx* Z = dalloc<cplx>(c, (size_t)b * x * d * w);
        std::vector<GemmItem> g2{sub_gemm(Y, GEMM_N, b * x, h->wr[i - 1], GEMM_N, d * v, Z, b * x, b * x, d * w,
                                          d * v, 0.0)};
        run_gemm(c, g2, prec);
        dfree(c, Y);
        return Z;
    }

    // Z[(c,a),(o,v)] for site i from L[i-1] (left environment of site i)
    cplx* z_from_left(int i) {
        const int b = cb(i - 1), cc = cb(i), a = ca(i - 1), v = dw(i), w = dw(i - 1);
        const cplx* A = psi->site[i - 1];
        const cplx* Li = L[i - 1];
        cplx* P = dalloc<cplx>(c, (size_t)cc * a * d * w);
        std::vector<GemmItem> g;
        for (int ww = 0; ww < w; ++ww)
            for (int ii = 0; ii < d; ++ii)
                g.push_back(sub_gemm(A + (size_t)b * ii, GEMM_T, b * d, Li + (size_t)b * a * ww, GEMM_N, b,
                                     P + (size_t)cc * a * (ii + d * ww), cc, cc, a, b, 0.0));
        run_gemm(c, g, prec);
        cplx* Q = dalloc<cplx>(c, (size_t)cc * a * d * v);
        std::vector<GemmItem> g2{sub_gemm(P, GEMM_N, cc * a, h->wl[i - 1], GEMM_N, d * w, Q, cc * a, cc * a, d * v,
                                          d * w, 0.0)};
        run_gemm(c, g2, prec);
        dfree(c, P);
        return Q;
    }

    // R[i-1] from R[i] and φ site i
    void right_step(int i) {
        const int b = cb(i - 1), x = ca(i), a = ca(i - 1), w = dw(i - 1);
        cplx* Z = z_from_right(i);
        const cplx* B = phi->site[i - 1];
        cplx* Rn = dalloc<cplx>(c, (size_t)b * a * w);
        for (int o = 0; o < d; ++o) {
            std::vector<GemmItem> g;
            for (int ww = 0; ww < w; ++ww)
                g.push_back(sub_gemm(Z + (size_t)b * x * (o + d * ww), GEMM_N, b, B + (size_t)a * o, GEMM_C, a * d,
                                     Rn + (size_t)b * a * ww, b, b, a, x, o == 0 ? 0.0 : 1.0));
            run_gemm(c, g, prec);
        }
        dfree(c, Z);
        if (R[i - 1]) dfree(c, R[i - 1]);
        R[i - 1] = Rn;
    }

    // L[i] from L[i-1] and φ site i
    void left_step(int i) {
        const int cc = cb(i), a = ca(i - 1), x = ca(i), v = dw(i);
        cplx* Q = z_from_left(i);
        const cplx* B = phi->site[i - 1];
        cplx* Ln = dalloc<cplx>(c, (size_t)cc * x * v);
        for (int o = 0; o < d; ++o) {
            std::vector<GemmItem> g;
            for (int vv = 0; vv < v; ++vv)
                g.push_back(sub_gemm(Q + (size_t)cc * a * (o + d * vv), GEMM_N, cc, B + (size_t)a * o, GEMM_CONJ,
                                     a * d, Ln + (size_t)cc * x * vv, cc, cc, x, a, o == 0 ? 0.0 : 1.0));
            run_gemm(c, g, prec);
        }
        dfree(c, Q);
        if (L[i]) dfree(c, L[i]);
        L[i] = Ln;
    }

    // the local tensor M[a, o, x] = projection of H|ψ⟩ on site i
    cplx* local(int i) {
        const int b = cb(i - 1), a = ca(i - 1), x = ca(i), w = dw(i - 1);
        cplx* Z = z_from_right(i);
        const cplx* Li = L[i - 1];
        cplx* M = dalloc<cplx>(c, (size_t)a * d * x);
        for (int ww = 0; ww < w; ++ww) {
            std::vector<GemmItem> g;
            for (int o = 0; o < d; ++o)
                g.push_back(sub_gemm(Li + (size_t)b * a * ww, GEMM_T, b, Z + (size_t)b * x * (o + d * ww), GEMM_N, b,
                                     M + (size_t)a * o, a * d, a, x, b, ww == 0 ? 0.0 : 1.0));
            run_gemm(c, g, prec);
        }
        dfree(c, Z);
        return M;
    }
};

void mpo_fit_core(qh_context* c, const qh_mpo* h, const qh_mps* psi, qh_mps* phi, int sweeps, int prec,
                  double* overlap) {
    check_mpo(c, h);
    check_mps(c, psi);
    check_mps(c, phi);
    if (psi == phi) throw Error(QH_ERR_ARGUMENT, "the fitted MPS must be distinct from the input MPS");
    if (psi->periodic || phi->periodic || h->periodic)
        throw Error(QH_ERR_ARGUMENT, "MPO fitting is defined for open chains only");
    if (psi->n != h->n || phi->n != h->n) throw Error(QH_ERR_DIMENSION, "MPO and MPS lengths differ");
    if (psi->d != h->d || phi->d != h->d) throw Error(QH_ERR_DIMENSION, "physical dimensions differ");
    if (sweeps < 1) throw Error(QH_ERR_ARGUMENT, "sweeps must be positive");
    const int n = h->n, d = h->d;
    if (n < 2) throw Error(QH_ERR_ARGUMENT, "an MPO needs at least 2 sites");

    // ψ's Λ vectors enter through an absorbed copy
    std::unique_ptr<qh_mps> pc(new qh_mps(*psi));
    std::vector<cplx*> own;
    for (int b = 1; b <= n - 1; ++b) {
        if (!psi->lam[b - 1]) continue;
        const size_t ne = site_elems(psi, b + 1);
        cplx* s = dalloc<cplx>(c, ne);
        QH_HIP(hipMemcpyAsync(s, psi->site[b], ne * sizeof(cplx), hipMemcpyDeviceToDevice, c->stream));
        launch_scale_bond(s, psi->chi[b], d, psi->chi[b + 1], psi->lam[b - 1], 0, 0, 0.0, c->stream);
        pc->site[b] = s;
        own.push_back(s);
    }

    // φ: no Λ, right-orthonormal from site 2 on
    absorb_lambdas(c, phi);
    for (int i = n; i >= 2; --i) {
        const int cl = phi->chi[i - 1], cr = phi->chi[i];
        // B_i = R Q with Q right-orthonormal (Q^H = orthonormal basis of B_i^H's
        // columns); R = B_i Q^H goes into B_{i-1}
        if (d * cr < cl) throw Error(QH_ERR_DIMENSION, "bond dimension exceeds the right dimension of the site");
        cplx* qh_ = dalloc<cplx>(c, (size_t)cl * d * cr);
        launch_conj_transpose(phi->site[i - 1], cl, d * cr, qh_, c->stream);  // (o r) × l
        orthonormalize_columns(c, qh_, d * cr, cl);
        cplx* r = dalloc<cplx>(c, (size_t)cl * cl);
        std::vector<GemmItem> g{gemm_item(phi->site[i - 1], GEMM_N, cl, qh_, GEMM_N, d * cr, r, cl, cl, cl, d * cr)};
        run_gemm(c, g);
        cplx* q = dalloc<cplx>(c, (size_t)cl * d * cr);
        launch_conj_transpose(qh_, d * cr, cl, q, c->stream);
        replace_site(phi, i, q);
        const int cll = phi->chi[i - 2];
        cplx* nb = dalloc<cplx>(c, (size_t)cll * d * cl);
        std::vector<GemmItem> g2{gemm_item(phi->site[i - 2], GEMM_N, cll * d, r, GEMM_N, cl, nb, cll * d, cll * d, cl,
                                           cl)};
        run_gemm(c, g2);
        replace_site(phi, i - 1, nb);
        dfree(c, qh_);
        dfree(c, r);
    }

    Fit f{c, h, pc.get(), phi, prec, n, d, std::vector<cplx*>(n + 1, nullptr), std::vector<cplx*>(n + 1, nullptr)};
    f.L[0] = f.ones();
    f.R[n] = f.ones();
    for (int i = n; i >= 2; --i) f.right_step(i);

    cplx* M = nullptr;
    for (int s = 0; s < sweeps; ++s) {
        for (int i = 1; i <= n - 1; ++i) {
            M = f.local(i);
            left_orthonormalize(c, M, phi->chi[i - 1], d, phi->chi[i]);
            replace_site(phi, i, M);
            f.left_step(i);
        }
        for (int i = n; i >= 2; --i) {
            M = f.local(i);
            right_orthonormalize(c, M, phi->chi[i - 1], d, phi->chi[i]);
            replace_site(phi, i, M);
            f.right_step(i);
        }
        M = f.local(1);
        replace_site(phi, 1, M);
    }
    // ⟨φ|Hψ⟩ = ⟨M|M⟩ with the orthogonality centre on site 1
    cplx* dot = dalloc<cplx>(c, 1);
    launch_dot(phi->site[0], phi->site[0], site_elems(phi, 1), dot, c->stream);
    cplx hv;
    readback(c, &hv, dot, sizeof(cplx));
    overlap[0] = hv.x;
    overlap[1] = hv.y;
    dfree(c, dot);
    for (auto p : f.L) if (p) dfree(c, p);
    for (auto p : f.R) if (p) dfree(c, p);
    for (auto p : own) dfree(c, p);
}

}  // namespace

static int mpo_create_impl(qh_context* c, int n, int d, const int* bond_dims, const double* const* arrays,
                           bool periodic, qh_mpo** out) {
    if (!c || !out) return QH_ERR_ARGUMENT;
    qh_mpo* h = nullptr;
    int rc = api_guard(c, [&] {
        if (n < 2) throw Error(QH_ERR_ARGUMENT, "an MPO needs at least 2 sites");
        if (d < 1) throw Error(QH_ERR_ARGUMENT, "invalid physical dimension");
        if (!arrays || !bond_dims) throw Error(QH_ERR_ARGUMENT, "null argument");
        h = new qh_mpo();
        h->ctx = c;
        h->n = n;
        h->d = d;
        h->periodic = periodic;
        h->D.assign(n + 1, 1);
        for (int b = 1; b <= (periodic ? n : n - 1); ++b) {
            if (bond_dims[b - 1] < 1) throw Error(QH_ERR_DIMENSION, "bond dimensions must be positive");
            h->D[b] = bond_dims[b - 1];
        }
        if (periodic) h->D[0] = h->D[n];
        h->w.assign(n, nullptr);
        h->wr.assign(n, nullptr);
        h->wl.assign(n, nullptr);
        for (int s = 1; s <= n; ++s) {
            if (!arrays[s - 1]) throw Error(QH_ERR_ARGUMENT, "null site array");
            const int dl = h->D[s - 1], dr = h->D[s];
            const zc* src = reinterpret_cast<const zc*>(arrays[s - 1]);
            std::vector<zc> w((size_t)dl * d * d * dr), wr(w.size()), wl(w.size());
            // reference (o, i, l, r) column-major -> W[w, o, i, v]
            for (int v = 0; v < dr; ++v)
                for (int ww = 0; ww < dl; ++ww)
                    for (int ii = 0; ii < d; ++ii)
                        for (int o = 0; o < d; ++o) {
                            const zc val = src[o + (size_t)d * (ii + (size_t)d * (ww + (size_t)dl * v))];
                            w[ww + (size_t)dl * (o + (size_t)d * (ii + (size_t)d * v))] = val;
                            wr[(ii + (size_t)d * v) + (size_t)(d * dr) * (o + (size_t)d * ww)] = val;
                            wl[(ii + (size_t)d * ww) + (size_t)(d * dl) * (o + (size_t)d * v)] = val;
                        }
            auto put = [&](const std::vector<zc>& hv) {
                cplx* p = dalloc<cplx>(c, hv.size());
                keep(c, p);
                QH_HIP(hipMemcpyAsync(p, hv.data(), hv.size() * sizeof(zc), hipMemcpyHostToDevice, c->stream));
                return p;
            };
            h->w[s - 1] = put(w);
            h->wr[s - 1] = put(wr);
            h->wl[s - 1] = put(wl);
        }
        ctx_sync(c);
    });
    if (rc != QH_OK) {
        if (h) {
            for (auto p : h->w) dfree(c, p);
            for (auto p : h->wr) dfree(c, p);
            for (auto p : h->wl) dfree(c, p);
            delete h;
        }
        return rc;
    }
    *out = h;
    return QH_OK;
}

extern "C" {

int qh_mpo_create(qh_context* c, int n, int d, const int* bond_dims, const double* const* arrays, qh_mpo** out) {
    return mpo_create_impl(c, n, d, bond_dims, arrays, false, out);
}

int qh_pmpo_create(qh_context* c, int n, int d, const int* bond_dims, const double* const* arrays, qh_mpo** out) {
    return mpo_create_impl(c, n, d, bond_dims, arrays, true, out);
}

int qh_mpo_free(qh_mpo* h) {
    if (!h) return QH_OK;
    qh_context* c = h->ctx;
    int rc = api_guard(c, [&] {
        for (auto p : h->w) dfree(c, p);
        for (auto p : h->wr) dfree(c, p);
        for (auto p : h->wl) dfree(c, p);
    });
    delete h;
    return rc;
}

int qh_mpo_apply(qh_context* c, const qh_mpo* h, const qh_mps* psi, qh_mps** out) {
    if (!c || !out) return QH_ERR_ARGUMENT;
    qh_mps* m = nullptr;
    int rc = api_guard(c, [&] {
        check_mpo(c, h);
        check_mps(c, psi);
        if (psi->n != h->n) throw Error(QH_ERR_DIMENSION, "MPO and MPS lengths differ");
        if (psi->d != h->d) throw Error(QH_ERR_DIMENSION, "physical dimensions differ");
        if (psi->periodic != h->periodic)
            throw Error(QH_ERR_ARGUMENT, "MPO and MPS must have the same boundary (both open or both periodic)");
        const int n = h->n, d = h->d;
        m = new qh_mps();
        m->ctx = c;
        m->n = n;
        m->d = d;
        m->periodic = psi->periodic;
        m->chi.assign(n + 1, 1);
        for (int b = 0; b <= n; ++b) m->chi[b] = psi->chi[b] * h->D[b];
        m->site.assign(n, nullptr);
        m->lam.assign(n - 1, nullptr);
        std::vector<cplx*> tmp;
        for (int s = 1; s <= n; ++s) {
            // ψ site with its right Λ absorbed
            const cplx* a = psi->site[s - 1];
            if (s <= n - 1 && psi->lam[s - 1]) {
                cplx* t = dalloc<cplx>(c, site_elems(psi, s));
                QH_HIP(hipMemcpyAsync(t, a, site_elems(psi, s) * sizeof(cplx), hipMemcpyDeviceToDevice, c->stream));
                launch_scale_bond(t, psi->chi[s - 1], d, psi->chi[s], psi->lam[s - 1], 2, 0, 0.0, c->stream);
                tmp.push_back(t);
                a = t;
            }
            m->site[s - 1] = dalloc<cplx>(c, site_elems(m, s));
            keep(c, m->site[s - 1]);
            launch_mpo_apply(h->w[s - 1], a, m->site[s - 1], psi->chi[s - 1], psi->chi[s], h->D[s - 1], h->D[s], d,
                             c->stream);
        }
        for (auto t : tmp) dfree(c, t);
        ctx_sync(c);
    });
    if (rc != QH_OK) {
        if (m) {
            for (auto p : m->site) dfree(c, p);
            delete m;
        }
        return rc;
    }
    *out = m;
    return QH_OK;
}

int qh_mpo_fit(qh_context* c, const qh_mpo* h, const qh_mps* psi, qh_mps* phi, int sweeps, unsigned flags,
               double* overlap) {
    if (!c || !overlap) return QH_ERR_ARGUMENT;
    return api_guard(c, [&] {
        const unsigned prec = flags & QH_PREC_MASK;
        if (prec != QH_PREC_FP64 && prec != QH_PREC_FP32 && prec != QH_PREC_BF16X3)
            throw Error(QH_ERR_ARGUMENT, "unknown contraction precision");
        mpo_fit_core(c, h, psi, phi, sweeps, (int)prec, overlap);
    });
}

}  // extern "C"
