// C ABI of libqrochet_hip.so: Qrochet's MPS update path on one MI355X.
// Each entry point restates one method of src/Ansatz/Chain.jl (cited per
// function) on device-resident tensors; see include/qrochet_hip.h.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <utility>
#include <vector>

#include "qh_runtime.h"

using namespace qh;

namespace qh {
void launch_apply_op(cplx* t, int left, int dim, int right, const cplx* op, hipStream_t s);
void launch_eye_rect(cplx* p, int m, int n, hipStream_t s);
void launch_trace(const cplx* e, int n, cplx* out, hipStream_t s);
}

namespace {

template <class F>
int guard(qh_context* c, F&& f) {
    return api_guard(c, std::forward<F>(f));
}

using zc = std::complex<double>;

void check_site(const qh_mps* m, int site) {
    if (site < 1 || site > m->n) throw Error(QH_ERR_ASSERTION, "Site " + std::to_string(site) + " not found");
}

void check_bond(const qh_mps* m, int bond) {
    if (bond < 1 || bond > m->n - 1) throw Error(QH_ERR_ARGUMENT, "Invalid bond (" + std::to_string(bond) + ", " +
                                                                       std::to_string(bond + 1) + ")");
}

// canonize_site!/canonize!/mixed_canonize! (and normalize!) dispatch on `::Open`
// only (src/Ansatz/Chain.jl:336,339,461,469,500,509)
void check_open(const qh_mps* m, const char* what) {
    if (m->periodic) throw Error(QH_ERR_ARGUMENT, std::string(what) + " is defined for open chains only");
}

cplx* dcopy(qh_context* c, const cplx* src, size_t n) {
    cplx* p = dalloc<cplx>(c, n);
    if (n) QH_HIP(hipMemcpyAsync(p, src, n * sizeof(cplx), hipMemcpyDeviceToDevice, c->stream));
    return p;
}
double* dcopy_d(qh_context* c, const double* src, size_t n) {
    double* p = dalloc<double>(c, n);
    if (n) QH_HIP(hipMemcpyAsync(p, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return p;
}

// ---------------------------------------------------------------------------
// evolve! for 2-site gates, batched over independent bonds
// (Chain.jl:606-661 with contract_2sitewf! :669-685 and unpack_2sitewf! :693-722)
// ---------------------------------------------------------------------------
void normalize_core(qh_context* c, qh_mps* m, int root);

void evolve2_core(qh_context* c, int count, qh_mps* const* ms, const int* bonds, const double* gates,
                  const double* thresholds, const int* maxdims, unsigned flags, int* kept_out, double* disc_out) {
    if (count <= 0) return;
    if (!ms || !bonds || !gates) throw Error(QH_ERR_ARGUMENT, "null argument");
    const unsigned prec = flags & QH_PREC_MASK;
    if (prec != QH_PREC_FP64 && prec != QH_PREC_FP32 && prec != QH_PREC_BF16X3)
        throw Error(QH_ERR_ARGUMENT, "unknown contraction precision");
    const bool canonical = flags & QH_CANONICAL;
    const bool renorm = flags & QH_RENORMALIZE;
    auto thr_of = [&](int i) { return thresholds ? thresholds[i] : std::nan(""); };
    auto md_of = [&](int i) { return maxdims ? maxdims[i] : -1; };
    for (int i = 0; i < count; ++i)
        if (md_of(i) == 0) throw Error(QH_ERR_ARGUMENT, "maxdim must be positive");
    const int d = ms[0] ? ms[0]->d : 0;
    for (int i = 0; i < count; ++i) {
        check_mps(c, ms[i]);
        check_bond(ms[i], bonds[i]);
        if (ms[i]->d != d) throw Error(QH_ERR_ARGUMENT, "all MPS of a batch must share the physical dimension");
        for (int j = 0; j < i; ++j)
            if (ms[j] == ms[i] && std::abs(bonds[j] - bonds[i]) < 2)
                throw Error(QH_ERR_ARGUMENT, "updates of one batch must act on disjoint sites");
    }
    const int d4 = d * d * d * d;
    std::vector<cplx> hg((size_t)count * d4);
    std::memcpy(hg.data(), gates, hg.size() * sizeof(cplx));
    const cplx* dg = upload(c, hg);

    std::vector<ThetaItem> th(count);
    std::vector<SvdTask> tasks(count);
    std::vector<cplx*> xbuf(count), nl(count), nr(count);
    std::vector<double*> nlam(count), pl(count, nullptr), pr(count, nullptr);
    for (int i = 0; i < count; ++i) {
        qh_mps* m = ms[i];
        const int b = bonds[i];
        const int chil = m->chi[b - 1], chim = m->chi[b], chir = m->chi[b + 1];
        const int mr = d * chil, mc = d * chir;
        xbuf[i] = dalloc<cplx>(c, (size_t)mr * mc);
        ThetaItem& t = th[i];
        t = ThetaItem{};
        t.al = m->site[b - 1];
        t.ar = m->site[b];
        t.lam = m->lam[b - 1];
        t.rs = (canonical && b >= 2) ? m->lam[b - 2] : nullptr;
        t.cs = (canonical && b + 1 <= m->n - 1) ? m->lam[b] : nullptr;
        t.gate = dg + (size_t)i * d4;
        t.x = xbuf[i];
        t.chil = chil; t.chim = chim; t.chir = chir; t.d = d;
        t.transposed = mr < mc ? 1 : 0;
        if (t.rs) { pl[i] = dalloc<double>(c, chil); launch_pinv(t.rs, chil, 1e-32, pl[i], c->stream); }
        if (t.cs) { pr[i] = dalloc<double>(c, chir); launch_pinv(t.cs, chir, 1e-32, pr[i], c->stream); }
        const double threshold = thr_of(i);
        const int maxdim = md_of(i);
        const bool trunc = !std::isnan(threshold) || maxdim >= 0;
        SvdTask& s = tasks[i];
        s.x = xbuf[i];
        s.mr = mr; s.mc = mc;
        s.truncate = trunc ? 1 : 0;
        s.threshold = std::isnan(threshold) ? 1e-16 : threshold;
        s.maxdim = maxdim;
        s.renormalize = (canonical && renorm && trunc) ? 1 : 0;
        s.refine = canonical ? 1 : 0;  // Vidal form: the factors are the canonical tensors
        const int kmax = s.kmax();
        nl[i] = dalloc<cplx>(c, (size_t)chil * d * kmax);
        nr[i] = dalloc<cplx>(c, (size_t)kmax * d * chir);
        nlam[i] = dalloc<double>(c, kmax);
        SvdOut u;
        u.factor = 0; u.conj_transpose = 0; u.rowscale = pl[i]; u.rowscale_mod = chil; u.rowscale_div = 1;
        u.dst = nl[i];
        SvdOut v;
        v.factor = 1; v.conj_transpose = 1; v.rowscale = pr[i]; v.rowscale_mod = chir; v.rowscale_div = d;
        v.dst = nr[i];
        s.outs = {u, v};
        s.lam_dst = nlam[i];
    }
    run_theta(c, th, (int)prec);
    run_svd_tasks(c, tasks);
    for (int i = 0; i < count; ++i) {
        qh_mps* m = ms[i];
        const int b = bonds[i];
        replace_site(m, b, nl[i]);
        replace_site(m, b + 1, nr[i]);
        replace_lam(m, b, nlam[i]);
        m->chi[b] = tasks[i].kept;
        dfree(c, xbuf[i]);
        dfree(c, pl[i]);
        dfree(c, pr[i]);
        if (kept_out) kept_out[i] = tasks[i].kept;
        if (disc_out) disc_out[i] = tasks[i].discarded;
    }
    if (renorm && !canonical)
        for (int i = 0; i < count; ++i)
            if (tasks[i].truncate) normalize_core(c, ms[i], bonds[i]);
}

// ---------------------------------------------------------------------------
// factorisations of one site (Chain.jl:339-376)
// ---------------------------------------------------------------------------
// `canonize_site!(…; method = :qr)` (Chain.jl:339-376) with a Householder QR like
// LAPACK's `qr!`: RIGHT: A (χl d × χr) = Q R, Q -> site, R Λ A_{i+1} -> site+1;
// LEFT: A (χl × d χr) = R^H Q^H from the QR of A^H, Q^H -> site, A_{i-1} Λ R^H -> site-1.
// The bond dimension stays χr (RIGHT) / χl (LEFT); the bond's Λ is absorbed.
void householder_site(qh_context* c, qh_mps* m, int site, int direction, int mr, int mc) {
    const int d = m->d;
    const int rows = direction == QH_RIGHT ? mr : mc;  // the tall matrix factored
    const int cols = direction == QH_RIGHT ? mc : mr;
    const cplx* a = m->site[site - 1];
    cplx* x = nullptr;
    if (direction == QH_RIGHT) x = dcopy(c, a, (size_t)mr * mc);
    else {
        x = dalloc<cplx>(c, (size_t)mr * mc);
        launch_conj_transpose(a, mr, mc, x, c->stream);  // A^H: (d χr) × χl
    }
    const size_t tq = (size_t)((cols + QR_NB - 1) / QR_NB) * QR_NB * QR_NB;
    cplx* v = dalloc<cplx>(c, (size_t)rows * cols);
    cplx* t = dalloc<cplx>(c, tq);
    std::vector<QrJob> jobs{QrJob{x, rows, cols, rows, v, t}};
    run_qr_factor(c, jobs);  // R in the upper triangle of x (ld rows), zeros below
    cplx* q = dalloc<cplx>(c, (size_t)rows * cols);
    launch_eye_rect(q, rows, cols, c->stream);
    std::vector<QrApplyJob> app{QrApplyJob{&jobs[0], q, cols}};
    run_qr_apply_q(c, app);
    if (direction == QH_RIGHT) {
        const int nb = site + 1;
        const int n2 = d * m->chi[nb];
        cplx* nn = dalloc<cplx>(c, (size_t)cols * n2);
        std::vector<GemmItem> g{gemm_item(x, GEMM_N, rows, m->site[nb - 1], GEMM_N, cols, nn, cols, cols, n2, cols)};
        g[0].dk = m->lam[site - 1];
        g[0].dk_mod = cols;
        run_gemm(c, g);
        replace_site(m, site, q);
        replace_site(m, nb, nn);
        replace_lam(m, site, nullptr);
    } else {
        const int pb = site - 1;
        const int m2 = m->chi[pb - 1] * d;
        cplx* na = dalloc<cplx>(c, (size_t)cols * rows);
        launch_conj_transpose(q, rows, cols, na, c->stream);  // Q^H: χl × (d χr)
        cplx* nn = dalloc<cplx>(c, (size_t)m2 * cols);
        std::vector<GemmItem> g{gemm_item(m->site[pb - 1], GEMM_N, m2, x, GEMM_C, rows, nn, m2, m2, cols, cols)};
        g[0].dk = m->lam[site - 2];
        g[0].dk_mod = cols;
        run_gemm(c, g);
        dfree(c, q);
        replace_site(m, site, na);
        replace_site(m, pb, nn);
        replace_lam(m, pb, nullptr);
    }
    dfree(c, x);
    dfree(c, v);
    dfree(c, t);
}

void canonize_site_core(qh_context* c, qh_mps* m, int site, int direction, int method) {
    check_open(m, "canonize_site!");
    check_site(m, site);
    if (direction == QH_LEFT) {
        if (site == 1) throw Error(QH_ERR_ARGUMENT, "Cannot right-canonize left-most tensor");
    } else if (direction == QH_RIGHT) {
        if (site == m->n) throw Error(QH_ERR_ARGUMENT, "Cannot left-canonize right-most tensor");
    } else {
        throw Error(QH_ERR_ARGUMENT, "Unknown direction");
    }
    if (method != QH_METHOD_QR && method != QH_METHOD_SVD) throw Error(QH_ERR_ARGUMENT, "Unknown factorization method");
    const int d = m->d;
    const int chil = m->chi[site - 1], chir = m->chi[site];
    // logical matrix: RIGHT -> A as (χl d) × χr ;  LEFT -> A as χl × (d χr)
    const int mr = direction == QH_RIGHT ? chil * d : chil;
    const int mc = direction == QH_RIGHT ? chir : d * chir;
    const int k = std::min(mr, mc);
    const cplx* a = m->site[site - 1];
    if (method == QH_METHOD_QR && ((direction == QH_RIGHT && mr >= mc) || (direction == QH_LEFT && mc >= mr))) {
        householder_site(c, m, site, direction, mr, mc);
        return;
    }
    cplx* x = nullptr;
    if (mr >= mc) x = dcopy(c, a, (size_t)mr * mc);
    else {
        x = dalloc<cplx>(c, (size_t)mr * mc);
        launch_conj_transpose(a, mr, mc, x, c->stream);
    }
    double* sig = dalloc<double>(c, k);
    SvdTask t;
    t.x = x; t.mr = mr; t.mc = mc;
    t.truncate = 0;
    t.refine = 1;  // the site must pass isleft/isrightcanonical like LAPACK's factors
    t.lam_dst = sig;
    cplx* na = nullptr;   // new tensor at site
    cplx* f2 = nullptr;   // factor passed to the neighbour
    if (direction == QH_RIGHT) {
        na = dalloc<cplx>(c, (size_t)mr * k);   // U (χl d × k)
        f2 = dalloc<cplx>(c, (size_t)k * mc);   // V^H (k × χr)
        SvdOut u; u.factor = 0; u.dst = na;
        SvdOut v; v.factor = 1; v.conj_transpose = 1; v.dst = f2;
        t.outs = {u, v};
    } else {
        na = dalloc<cplx>(c, (size_t)k * mc);   // V^H (k × d χr)
        f2 = dalloc<cplx>(c, (size_t)mr * k);   // U (χl × k)
        SvdOut v; v.factor = 1; v.conj_transpose = 1; v.dst = na;
        SvdOut u; u.factor = 0; u.dst = f2;
        t.outs = {v, u};
    }
    std::vector<SvdTask> tasks{t};
    run_svd_tasks(c, tasks);
    dfree(c, x);
    if (direction == QH_RIGHT) {
        // A_{i+1} <- (S) V^H · Λ_old · A_{i+1}
        const int nb = site + 1;
        const int n2 = d * m->chi[nb];
        cplx* nn = dalloc<cplx>(c, (size_t)k * n2);
        std::vector<GemmItem> g{gemm_item(f2, 0, k, m->site[nb - 1], 0, chir, nn, k, k, n2, chir)};
        g[0].dk = m->lam[site - 1];
        g[0].dk_mod = chir;
        if (method == QH_METHOD_QR) g[0].rowscale = sig;
        run_gemm(c, g);
        replace_site(m, site, na);
        replace_site(m, nb, nn);
        if (method == QH_METHOD_SVD) replace_lam(m, site, sig);
        else { replace_lam(m, site, nullptr); dfree(c, sig); }
        m->chi[site] = k;
    } else {
        // A_{i-1} <- A_{i-1} · Λ_old · U (S)
        const int pb = site - 1;
        const int m2 = m->chi[pb - 1] * d;
        cplx* nn = dalloc<cplx>(c, (size_t)m2 * k);
        std::vector<GemmItem> g{gemm_item(m->site[pb - 1], 0, m2, f2, 0, chil, nn, m2, m2, k, chil)};
        g[0].dk = m->lam[site - 2];
        g[0].dk_mod = chil;
        if (method == QH_METHOD_QR) g[0].colscale = sig;
        run_gemm(c, g);
        replace_site(m, site, na);
        replace_site(m, pb, nn);
        if (method == QH_METHOD_SVD) replace_lam(m, pb, sig);
        else { replace_lam(m, pb, nullptr); dfree(c, sig); }
        m->chi[site - 1] = k;
    }
    dfree(c, f2);
}

// Chain.jl:390-422
void truncate_core(qh_context* c, qh_mps* m, int bond, double threshold, int maxdim) {
    check_bond(m, bond);
    if (!m->lam[bond - 1]) throw Error(QH_ERR_MISSING_SCHMIDT, "Can't access the spectrum on bond (" +
                                                                   std::to_string(bond) + ", " +
                                                                   std::to_string(bond + 1) + ")");
    if (maxdim == 0) throw Error(QH_ERR_ARGUMENT, "maxdim must be positive");
    const int chi = m->chi[bond];
    std::vector<double> lam(chi);
    readback(c, lam.data(), m->lam[bond - 1], chi * sizeof(double));
    const int lim = maxdim >= 0 ? std::min(chi, maxdim) : chi;
    const double thr = std::isnan(threshold) ? 1e-16 : threshold;
    std::vector<int> keep;
    std::vector<double> nl;
    for (int i = 0; i < lim; ++i)
        if (std::fabs(lam[i]) > thr) { keep.push_back(i); nl.push_back(lam[i]); }
    const int kp = (int)keep.size();
    if (kp == chi) return;
    const int d = m->d;
    const int chil = m->chi[bond - 1], chir = m->chi[bond + 1];
    std::vector<int> hk{kp};
    const int* dperm = upload(c, keep.empty() ? std::vector<int>{0} : keep);
    const int* dkept = upload(c, hk);
    cplx* al = dalloc<cplx>(c, (size_t)chil * d * kp);
    cplx* ar = dalloc<cplx>(c, (size_t)kp * d * chir);
    cplx* tmp = dalloc<cplx>(c, (size_t)chi * d * chir);
    launch_conj_transpose(m->site[bond], chi, d * chir, tmp, c->stream);  // (d χr) × χ
    std::vector<GatherItem> g(2);
    g[0] = GatherItem{};
    g[0].src = m->site[bond - 1]; g[0].ld_src = chil * d; g[0].rows = chil * d;
    g[0].perm = dperm; g[0].kept = dkept; g[0].out = al; g[0].max_cols = kp;
    g[0].rowscale_mod = 1; g[0].rowscale_div = 1;
    g[1] = GatherItem{};
    g[1].src = tmp; g[1].ld_src = d * chir; g[1].rows = d * chir;
    g[1].perm = dperm; g[1].kept = dkept; g[1].out = ar; g[1].conj_transpose = 1; g[1].max_cols = kp;
    g[1].rowscale_mod = 1; g[1].rowscale_div = 1;
    run_gather(c, g);
    double* dl = dalloc<double>(c, kp);
    if (kp) {
        const double* src = upload(c, nl);
        QH_HIP(hipMemcpyAsync(dl, src, kp * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    }
    replace_site(m, bond, al);
    replace_site(m, bond + 1, ar);
    replace_lam(m, bond, dl);
    m->chi[bond] = kp;
    dfree(c, tmp);
    ctx_sync(c);
}

// Chain.jl:309-333
void contract_between_core(qh_context* c, qh_mps* m, int s1, int s2, int direction, int del) {
    check_site(m, s1);
    check_site(m, s2);
    if (s1 == s2) throw Error(QH_ERR_ASSERTION, "Sites must be different");
    if (std::abs(s1 - s2) != 1) return;
    const int bond = std::min(s1, s2);
    double* lam = m->lam[bond - 1];
    if (!lam) return;
    int target, mode;
    if (direction == QH_RIGHT) { target = s2; mode = s2 > s1 ? 0 : 2; }
    else if (direction == QH_LEFT) { target = s1; mode = s2 > s1 ? 2 : 0; }
    else throw Error(QH_ERR_ARGUMENT, "Unknown direction");
    launch_scale_bond(m->site[target - 1], m->chi[target - 1], m->d, m->chi[target], lam, mode, 0, 0.0, c->stream);
    if (del) replace_lam(m, bond, nullptr);
}

// Chain.jl:469-497
// The left-to-right half of canonize! (Chain.jl:469-497) after the right-to-left QR
// sweep, as the reference takes it: n-1 dependent SVDs, each absorbing the previous
// one's s V^H.  Used when a bond is wider than d times its left neighbour.
void canonize_svd_sweep(qh_context* c, qh_mps* m) {
    const int n = m->n;
    std::vector<double*> lams(n - 1, nullptr);
    for (int i = 1; i <= n - 1; ++i) {
        canonize_site_core(c, m, i, QH_RIGHT, QH_METHOD_SVD);
        lams[i - 1] = m->lam[i - 1];
        m->lam[i - 1] = nullptr;
        launch_scale_bond(m->site[i], m->chi[i], m->d, m->chi[i + 1], lams[i - 1], 0, 0, 0.0, c->stream);
    }
    for (int i = 2; i <= n; ++i) {
        launch_scale_bond(m->site[i - 1], m->chi[i - 1], m->d, m->chi[i], lams[i - 2], 0, 1, 1e-64, c->stream);
        m->lam[i - 2] = lams[i - 2];
    }
    ctx_sync(c);
}

// `canonize!(ψ)` (Chain.jl:469-497): Vidal form Γ_1 Λ_1 Γ_2 … Γ_n with the Schmidt
// values on every bond.  After the reference's right-to-left QR sweep (sites 2..n
// right isometries B_i), a left-to-right sweep of QRs M_i = Q_i R_i with
// M_1 = A_1, M_i = R_{i-1} B_i gives ψ = Q_1⋯Q_i · R_i · B_{i+1}⋯B_n for every bond i,
// an isometry on each side of R_i: the SVDs R_i = U_i s_i V_i^H are the Schmidt
// decompositions the reference's dependent SVD sweep finds one after the other, and
// here they are independent (one batched call).  Then Λ_i = s_i and
//   Γ_1 = Q_1 U_1,   Γ_i = pinv(s_{i-1}) U_{i-1}^H Q_i U_i,   Γ_n = pinv(s_{n-1}) U_{n-1}^H M_n
// (pinv with atol 1e-64, Chain.jl:491) -- the reference's tensors up to the gauge of
// each singular subspace.  The QR sweep needs every M_i tall (χ_{i-1} d >= χ_i);
// otherwise the reference's SVD sweep runs instead.
void canonize_core(qh_context* c, qh_mps* m) {
    check_open(m, "canonize!");
    const int n = m->n, d = m->d;
    for (int i = n; i >= 2; --i) canonize_site_core(c, m, i, QH_LEFT, QH_METHOD_QR);
    for (int i = 1; i <= n - 1; ++i)
        if (m->chi[i - 1] * d < m->chi[i]) {
            canonize_svd_sweep(c, m);
            return;
        }
    // left-to-right QR sweep: Q_i (χ_{i-1} d × χ_i), R_i^H (χ_i × χ_i)
    std::vector<cplx*> Q(n + 1, nullptr), RH(n + 1, nullptr);
    cplx* M = dcopy(c, m->site[0], site_elems(m, 1));
    for (int i = 1; i <= n - 1; ++i) {
        const int rows = m->chi[i - 1] * d, cols = m->chi[i];
        const size_t tq = (size_t)((cols + QR_NB - 1) / QR_NB) * QR_NB * QR_NB;
        cplx* v = dalloc<cplx>(c, (size_t)rows * cols);
        cplx* t = dalloc<cplx>(c, tq);
        std::vector<QrJob> jobs{QrJob{M, rows, cols, rows, v, t}};
        run_qr_factor(c, jobs);
        Q[i] = dalloc<cplx>(c, (size_t)rows * cols);
        launch_eye_rect(Q[i], rows, cols, c->stream);
        std::vector<QrApplyJob> app{QrApplyJob{&jobs[0], Q[i], cols}};
        run_qr_apply_q(c, app);
        RH[i] = dalloc<cplx>(c, (size_t)cols * cols);
        std::vector<UpperHItem> uh{UpperHItem{M, rows, cols, RH[i]}};
        std::vector<int2> ut;
        for (size_t b = 0; b * 256 < (size_t)cols * cols; ++b) ut.push_back(make_int2(0, (int)b));
        UpperHItem* duh;
        int2* dut;
        upload2(c, uh, ut, duh, dut);
        launch_upper_h_batch(duh, dut, (int)ut.size(), c->stream);
        // M_{i+1} = R_i B_{i+1}  (χ_i × d χ_{i+1}); R_i = (R_i^H)^H
        const int n2 = d * m->chi[i + 1];
        cplx* nxt = dalloc<cplx>(c, (size_t)cols * n2);
        std::vector<GemmItem> g{gemm_item(RH[i], GEMM_C, cols, m->site[i], GEMM_N, cols, nxt, cols, cols, n2, cols)};
        run_gemm(c, g);
        dfree(c, M);
        dfree(c, v);
        dfree(c, t);
        M = nxt;
    }
    // SVDs of every bond matrix at once: the logical matrix is R_i^H = V s U^H, so the
    // factor wanted, U_i, is its right factor
    std::vector<SvdTask> tasks(n - 1);
    std::vector<cplx*> U(n + 1, nullptr);
    std::vector<double*> lam(n + 1, nullptr);
    for (int i = 1; i <= n - 1; ++i) {
        const int k = m->chi[i];
        U[i] = dalloc<cplx>(c, (size_t)k * k);
        lam[i] = dalloc<double>(c, k);
        SvdTask& t = tasks[i - 1];
        t.x = RH[i];
        t.mr = t.mc = k;
        t.truncate = 0;
        t.refine = 1;
        t.lam_dst = lam[i];
        SvdOut o;
        o.factor = 1;
        o.dst = U[i];
        t.outs = {o};
    }
    run_svd_tasks(c, tasks);
    // Γ_i = pinv(s_{i-1}) U_{i-1}^H (Q_i U_i), batched over the sites
    std::vector<double*> pinv(n + 1, nullptr);
    for (int i = 1; i <= n - 1; ++i) {
        pinv[i] = dalloc<double>(c, m->chi[i]);
        launch_pinv(lam[i], m->chi[i], 1e-64, pinv[i], c->stream);
    }
    std::vector<cplx*> QU(n + 1, nullptr), G(n + 1, nullptr);
    std::vector<GemmItem> g1, g2;
    for (int i = 1; i <= n - 1; ++i) {
        const int rows = m->chi[i - 1] * d, k = m->chi[i];
        QU[i] = dalloc<cplx>(c, (size_t)rows * k);
        g1.push_back(gemm_item(Q[i], GEMM_N, rows, U[i], GEMM_N, k, QU[i], rows, rows, k, k));
    }
    run_gemm(c, g1);
    G[1] = QU[1];
    for (int i = 2; i <= n; ++i) {
        const int kl = m->chi[i - 1];
        const int cols = d * m->chi[i];
        const cplx* src = (i == n) ? M : QU[i];  // (χ_{i-1} × d χ_i)
        G[i] = dalloc<cplx>(c, (size_t)kl * cols);
        GemmItem g = gemm_item(U[i - 1], GEMM_C, kl, src, GEMM_N, kl, G[i], kl, kl, cols, kl);
        g.rowscale = pinv[i - 1];
        g2.push_back(g);
    }
    run_gemm(c, g2);
    for (int i = 1; i <= n; ++i) replace_site(m, i, G[i]);
    for (int i = 1; i <= n - 1; ++i) replace_lam(m, i, lam[i]);
    for (int i = 2; i <= n - 1; ++i) dfree(c, QU[i]);
    for (int i = 1; i <= n - 1; ++i) {
        dfree(c, Q[i]);
        dfree(c, RH[i]);
        dfree(c, U[i]);
        dfree(c, pinv[i]);
    }
    dfree(c, M);
    ctx_sync(c);
}

// Chain.jl:509-524
void mixed_canonize_core(qh_context* c, qh_mps* m, int center) {
    check_open(m, "mixed_canonize!");
    check_site(m, center);
    for (int i = 1; i <= center - 1; ++i) canonize_site_core(c, m, i, QH_RIGHT, QH_METHOD_QR);
    for (int i = m->n; i >= center + 1; --i) canonize_site_core(c, m, i, QH_LEFT, QH_METHOD_QR);
    canonize_site_core(c, m, center, QH_LEFT, QH_METHOD_SVD);
}

// Chain.jl:532-536
void normalize_core(qh_context* c, qh_mps* m, int root) {
    mixed_canonize_core(c, m, root);
    const int bond = root - 1;
    const int chi = m->chi[bond];
    std::vector<double> lam(chi);
    readback(c, lam.data(), m->lam[bond - 1], chi * sizeof(double));
    double s = 0.0;
    for (double v : lam) s += v * v;
    s = std::sqrt(s);
    for (double& v : lam) v /= s;
    const double* src = upload(c, lam);
    QH_HIP(hipMemcpyAsync(m->lam[bond - 1], src, chi * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    ctx_sync(c);
}

// ---------------------------------------------------------------------------
// environments (full-network contraction of Chain.jl:724-748, Ansatz.jl:107-109)
// ---------------------------------------------------------------------------
struct Cluster {
    int start = 0, width = 0;
    std::vector<zc> op;  // d^w × d^w, column-major, site `start` fastest
};

// embed op on [s, s+w) into [S, S+W)
std::vector<zc> embed(const std::vector<zc>& op, int d, int s, int w, int S, int W) {
    int D = 1;
    for (int i = 0; i < W; ++i) D *= d;
    int dw = 1;
    for (int i = 0; i < w; ++i) dw *= d;
    int lo = 1;
    for (int i = 0; i < s - S; ++i) lo *= d;
    std::vector<zc> out((size_t)D * D, zc(0.0, 0.0));
    for (int mu = 0; mu < D; ++mu)
        for (int nu = 0; nu < D; ++nu) {
            const int mlo = mu % lo, nlo = nu % lo;
            const int mhi = mu / (lo * dw), nhi = nu / (lo * dw);
            if (mlo != nlo || mhi != nhi) continue;
            const int mm = (mu / lo) % dw, nn = (nu / lo) % dw;
            out[mu + (size_t)D * nu] = op[mm + (size_t)dw * nn];
        }
    return out;
}

std::vector<zc> matmul(const std::vector<zc>& a, const std::vector<zc>& b, int D) {
    std::vector<zc> out((size_t)D * D, zc(0.0, 0.0));
    for (int j = 0; j < D; ++j)
        for (int k = 0; k < D; ++k) {
            const zc bk = b[k + (size_t)D * j];
            if (bk == zc(0.0, 0.0)) continue;
            for (int i = 0; i < D; ++i) out[i + (size_t)D * j] += a[i + (size_t)D * k] * bk;
        }
    return out;
}

// compose an ordered observable list into disjoint clusters (later ops act after earlier ones)
std::vector<Cluster> make_clusters(const qh_mps* m, int nobs, const qh_observable* obs) {
    const int d = m->d;
    std::vector<Cluster> cl;
    for (int k = 0; k < nobs; ++k) {
        const qh_observable& o = obs[k];
        if (o.nsites != 1 && o.nsites != 2) throw Error(QH_ERR_ARGUMENT, "observables act on 1 or 2 sites");
        if (!o.op) throw Error(QH_ERR_ARGUMENT, "null observable");
        if (o.site < 1 || o.site + o.nsites - 1 > m->n)
            throw Error(QH_ERR_ARGUMENT, "Gate inputs must be a subset of the TN sites");
        int dw = o.nsites == 1 ? d : d * d;
        Cluster cur;
        cur.start = o.site;
        cur.width = o.nsites;
        cur.op.resize((size_t)dw * dw);
        std::memcpy(cur.op.data(), o.op, cur.op.size() * sizeof(zc));
        // merge with every overlapping cluster
        int S = cur.start, E = cur.start + cur.width - 1;
        std::vector<Cluster> keep, over;
        for (auto& c2 : cl) {
            const int s2 = c2.start, e2 = c2.start + c2.width - 1;
            if (e2 < S || s2 > E) keep.push_back(c2);
            else { over.push_back(c2); }
        }
        for (auto& c2 : over) { S = std::min(S, c2.start); E = std::max(E, c2.start + c2.width - 1); }
        const int W = E - S + 1;
        int D = 1;
        for (int i = 0; i < W; ++i) D *= d;
        if (W > 4 || D > 16) throw Error(QH_ERR_ARGUMENT, "observables overlapping over more than 4 sites");
        std::vector<zc> prev((size_t)D * D, zc(0.0, 0.0));
        for (int i = 0; i < D; ++i) prev[i + (size_t)D * i] = 1.0;
        for (auto& c2 : over) prev = matmul(embed(c2.op, d, c2.start, c2.width, S, W), prev, D);
        Cluster merged;
        merged.start = S;
        merged.width = W;
        merged.op = matmul(embed(cur.op, d, cur.start, cur.width, S, W), prev, D);
        keep.push_back(merged);
        cl.swap(keep);
    }
    std::sort(cl.begin(), cl.end(), [](const Cluster& a, const Cluster& b) { return a.start < b.start; });
    return cl;
}

int ipow(int b, int e) {
    int r = 1;
    for (int i = 0; i < e; ++i) r *= b;
    return r;
}

// Advance the left environment E (χ_b × χ_a on bond s-1, without that bond's Λ)
// through sites s..e with the given clusters; returns the environment on bond e.
cplx* env_segment(qh_context* c, const qh_mps* ket, const qh_mps* bra, cplx* E, int s, int e,
                  const std::vector<Cluster>& cls) {
    const int d = ket->d;
    size_t ci = 0;
    while (ci < cls.size() && cls[ci].start < s) ++ci;
    int i = s;
    while (i <= e) {
        const int lb = bra->chi[i - 1];
        const bool has_cluster = ci < cls.size() && cls[ci].start == i;
        const int w = has_cluster ? cls[ci].width : 1;
        // ket: T = E Λa_{i-1} A_i A_{i+1} ...
        cplx* T = dalloc<cplx>(c, (size_t)lb * d * ket->chi[i]);
        {
            std::vector<GemmItem> g{gemm_item(E, 0, lb, ket->site[i - 1], 0, ket->chi[i - 1], T, lb, lb, d * ket->chi[i],
                                              ket->chi[i - 1])};
            if (i >= 2) { g[0].dk = ket->lam[i - 2]; g[0].dk_mod = ket->chi[i - 1]; }
            run_gemm(c, g);
        }
        dfree(c, E);
        for (int s2 = 1; s2 < w; ++s2) {
            const int j = i + s2;
            const int rows = lb * ipow(d, s2);
            cplx* T2 = dalloc<cplx>(c, (size_t)rows * d * ket->chi[j]);
            std::vector<GemmItem> g{gemm_item(T, 0, rows, ket->site[j - 1], 0, ket->chi[j - 1], T2, rows, rows,
                                              d * ket->chi[j], ket->chi[j - 1])};
            g[0].dk = ket->lam[j - 2];
            g[0].dk_mod = ket->chi[j - 1];
            run_gemm(c, g);
            dfree(c, T);
            T = T2;
        }
        const int ra = ket->chi[i + w - 1];
        if (has_cluster) {
            std::vector<cplx> hop(cls[ci].op.size());
            std::memcpy(hop.data(), cls[ci].op.data(), hop.size() * sizeof(cplx));
            const cplx* dop = upload(c, hop);
            launch_apply_op(T, lb, ipow(d, w), ra, dop, c->stream);
        }
        // bra: peel one site at a time
        int curl = lb;
        for (int s2 = 0; s2 < w; ++s2) {
            const int j = i + s2;
            const int rest = ipow(d, w - 1 - s2) * ra;
            const int rb = bra->chi[j];
            cplx* W = dalloc<cplx>(c, (size_t)rb * rest);
            std::vector<GemmItem> g{gemm_item(bra->site[j - 1], 1, curl * d, T, 0, curl * d, W, rb, rb, rest, curl * d)};
            if (j >= 2) { g[0].dk = bra->lam[j - 2]; g[0].dk_mod = curl; }
            run_gemm(c, g);
            dfree(c, T);
            T = W;
            curl = rb;
        }
        E = T;
        i += w;
        if (has_cluster) ++ci;
    }
    return E;
}

cplx* unit_env(qh_context* c) {
    std::vector<cplx> one{make_double2(1.0, 0.0)};
    const cplx* src = upload(c, one);
    cplx* e = dalloc<cplx>(c, 1);
    QH_HIP(hipMemcpyAsync(e, src, sizeof(cplx), hipMemcpyDeviceToDevice, c->stream));
    return e;
}

// Periodic chains: the wrap bonds of bra (Wb) and ket (Wk) ride along as a passenger
// index p = (wb, wk), P = Wb·Wk; the environment E[b, a, p] starts as the identity
// E[(b, a), p] = δ((b, a), p) and every step is the open-chain step batched over p.
// The value is the trace Σ_p E[(b, a) = p, p] after the last site.  Cost
// O(n · Wb·Wk · χ³): meant for moderate bond dimensions.
zc periodic_contraction(qh_context* c, const qh_mps* ket, const qh_mps* bra, const std::vector<Cluster>& cls) {
    const int d = ket->d, n = ket->n;
    const int Wk = ket->chi[0], Wb = bra->chi[0];
    const int P = Wk * Wb;
    cplx* E = dalloc<cplx>(c, (size_t)P * P);
    launch_eye_rect(E, P, P, c->stream);
    size_t ci = 0;
    int i = 1;
    while (i <= n) {
        const int lb = bra->chi[i - 1], la = ket->chi[i - 1];
        const bool has_cluster = ci < cls.size() && cls[ci].start == i;
        const int w = has_cluster ? cls[ci].width : 1;
        const int ra1 = ket->chi[i];
        cplx* T = dalloc<cplx>(c, (size_t)lb * d * ra1 * P);
        {
            std::vector<GemmItem> g;
            for (int p = 0; p < P; ++p) {
                g.push_back(gemm_item(E + (size_t)lb * la * p, 0, lb, ket->site[i - 1], 0, la,
                                      T + (size_t)lb * d * ra1 * p, lb, lb, d * ra1, la));
                if (i >= 2) { g.back().dk = ket->lam[i - 2]; g.back().dk_mod = la; }
            }
            run_gemm(c, g);
        }
        dfree(c, E);
        for (int s2 = 1; s2 < w; ++s2) {
            const int j = i + s2;
            const int rows = lb * ipow(d, s2);
            const int kj = ket->chi[j - 1], rj = ket->chi[j];
            cplx* T2 = dalloc<cplx>(c, (size_t)rows * d * rj * P);
            std::vector<GemmItem> g;
            for (int p = 0; p < P; ++p) {
                g.push_back(gemm_item(T + (size_t)rows * kj * p, 0, rows, ket->site[j - 1], 0, kj,
                                      T2 + (size_t)rows * d * rj * p, rows, rows, d * rj, kj));
                g.back().dk = ket->lam[j - 2];
                g.back().dk_mod = kj;
            }
            run_gemm(c, g);
            dfree(c, T);
            T = T2;
        }
        const int ra = ket->chi[i + w - 1];
        if (has_cluster) {
            std::vector<cplx> hop(cls[ci].op.size());
            std::memcpy(hop.data(), cls[ci].op.data(), hop.size() * sizeof(cplx));
            const cplx* dop = upload(c, hop);
            launch_apply_op(T, lb, ipow(d, w), ra * P, dop, c->stream);
        }
        int curl = lb;
        for (int s2 = 0; s2 < w; ++s2) {
            const int j = i + s2;
            const int rest = ipow(d, w - 1 - s2) * ra;
            const int rb = bra->chi[j];
            cplx* Wm = dalloc<cplx>(c, (size_t)rb * rest * P);
            std::vector<GemmItem> g;
            for (int p = 0; p < P; ++p) {
                g.push_back(gemm_item(bra->site[j - 1], 1, curl * d, T + (size_t)curl * d * rest * p, 0, curl * d,
                                      Wm + (size_t)rb * rest * p, rb, rb, rest, curl * d));
                if (j >= 2) { g.back().dk = bra->lam[j - 2]; g.back().dk_mod = curl; }
            }
            run_gemm(c, g);
            dfree(c, T);
            T = Wm;
            curl = rb;
        }
        E = T;
        i += w;
        if (has_cluster) ++ci;
    }
    cplx* tr = dalloc<cplx>(c, 1);
    launch_trace(E, P, tr, c->stream);
    cplx h;
    readback(c, &h, tr, sizeof(cplx));
    dfree(c, tr);
    dfree(c, E);
    return zc(h.x, h.y);
}

zc full_contraction(qh_context* c, const qh_mps* ket, const qh_mps* bra, const std::vector<Cluster>& cls) {
    if (ket->periodic != bra->periodic) throw Error(QH_ERR_ARGUMENT, "cannot contract a periodic with an open chain");
    if (ket->periodic) return periodic_contraction(c, ket, bra, cls);
    cplx* E = env_segment(c, ket, bra, unit_env(c), 1, ket->n, cls);
    cplx h;
    readback(c, &h, E, sizeof(cplx));
    dfree(c, E);
    return zc(h.x, h.y);
}

// right environments R_i (ket × bra on bond i, without Λ_i), i = 0..n
std::vector<cplx*> right_envs(qh_context* c, const qh_mps* ket, const qh_mps* bra) {
    const int n = ket->n, d = ket->d;
    std::vector<cplx*> R(n + 1, nullptr);
    R[n] = unit_env(c);
    for (int i = n; i >= 1; --i) {
        const int la = ket->chi[i - 1], ra = ket->chi[i], lb = bra->chi[i - 1], rb = bra->chi[i];
        cplx* T = dalloc<cplx>(c, (size_t)la * d * rb);
        std::vector<GemmItem> g{gemm_item(ket->site[i - 1], 0, la * d, R[i], 0, ra, T, la * d, la * d, rb, ra)};
        if (i <= n - 1) { g[0].dk = ket->lam[i - 1]; g[0].dk_mod = ra; }
        run_gemm(c, g);
        R[i - 1] = dalloc<cplx>(c, (size_t)la * lb);
        std::vector<GemmItem> g2{gemm_item(T, 0, la, bra->site[i - 1], 1, lb, R[i - 1], la, la, lb, d * rb)};
        if (i <= n - 1) { g2[0].dk = bra->lam[i - 1]; g2[0].dk_mod = rb; g2[0].dk_div = d; }
        run_gemm(c, g2);
        dfree(c, T);
    }
    return R;
}

std::vector<cplx*> left_envs(qh_context* c, const qh_mps* ket, const qh_mps* bra) {
    const int n = ket->n;
    std::vector<cplx*> L(n + 1, nullptr);
    L[0] = unit_env(c);
    for (int i = 1; i <= n; ++i) {
        const size_t sz = (size_t)bra->chi[i - 1] * ket->chi[i - 1];
        cplx* E = dcopy(c, L[i - 1], sz);
        L[i] = env_segment(c, ket, bra, E, i, i, {});
    }
    return L;
}


// Left environments L_i (bond i, without Λ_i) and right environments R_i of one open
// chain (bra = ket), both sweeps advanced in lockstep: each launch carries the left
// sweep's step at site k and the right sweep's at site n + 1 - k (the GEMMs of
// env_segment and right_envs, so the values are the same).
void lr_envs(qh_context* c, const qh_mps* m, std::vector<cplx*>& L, std::vector<cplx*>& R) {
    const int n = m->n, d = m->d;
    L.assign(n + 1, nullptr);
    R.assign(n + 1, nullptr);
    L[0] = unit_env(c);
    R[n] = unit_env(c);
    for (int k = 1; k <= n; ++k) {
        const int i = k, j = n + 1 - k;  // left site, right site
        const int la = m->chi[i - 1], ra = m->chi[i];
        cplx* TL = dalloc<cplx>(c, (size_t)la * d * ra);
        const int ja = m->chi[j - 1], jr = m->chi[j];
        cplx* TR = dalloc<cplx>(c, (size_t)ja * d * jr);
        std::vector<GemmItem> g1{gemm_item(L[i - 1], GEMM_N, la, m->site[i - 1], GEMM_N, la, TL, la, la, d * ra, la),
                                 gemm_item(m->site[j - 1], GEMM_N, ja * d, R[j], GEMM_N, jr, TR, ja * d, ja * d, jr,
                                           jr)};
        if (i >= 2) { g1[0].dk = m->lam[i - 2]; g1[0].dk_mod = la; }
        if (j <= n - 1) { g1[1].dk = m->lam[j - 1]; g1[1].dk_mod = jr; }
        run_gemm(c, g1);
        L[i] = dalloc<cplx>(c, (size_t)ra * ra);
        R[j - 1] = dalloc<cplx>(c, (size_t)ja * ja);
        std::vector<GemmItem> g2{gemm_item(m->site[i - 1], GEMM_C, la * d, TL, GEMM_N, la * d, L[i], ra, ra, ra, la * d),
                                 gemm_item(TR, GEMM_N, ja, m->site[j - 1], GEMM_C, ja, R[j - 1], ja, ja, ja, d * jr)};
        if (i >= 2) { g2[0].dk = m->lam[i - 2]; g2[0].dk_mod = la; }
        if (j <= n - 1) { g2[1].dk = m->lam[j - 1]; g2[1].dk_mod = jr; g2[1].dk_div = d; }
        run_gemm(c, g2);
        dfree(c, TL);
        dfree(c, TR);
    }
}

// All terms of expect_many on one open chain, their environment segments advanced
// together.  A term is a sequence of site steps from its first to its last site
// (site i, and the cluster of operators starting there if any, as in env_segment);
// the terms form a prefix tree over (first site, steps), so terms that begin alike
// share their common steps.  The tree is processed depth by depth; within a depth
// the micro-operations of all nodes (absorb a ket site, apply the cluster, peel a
// bra site) go out as one batched GEMM or one batched operator launch per phase.
// vals[t] receives term t's value.
void expect_terms_batched(qh_context* c, const qh_mps* m, const std::vector<std::vector<Cluster>>& terms,
                          cplx* vals) {
    const int n = m->n, d = m->d;
    struct Node {
        int parent = -1, depth = 0;
        int start = 0;            // roots: first site of their terms
        int site = 0, width = 1;  // the site step (non-roots)
        const Cluster* cl = nullptr;
        cplx* E = nullptr;        // environment after the step (roots: L[start - 1], borrowed)
        bool leaf = false;
        int children = 0;
    };
    std::vector<Node> nodes;
    std::map<std::pair<int, std::string>, int> index;  // (parent or -start-1, step key) -> node
    std::vector<int> term_leaf(terms.size());
    std::vector<int> term_end(terms.size());
    int maxdepth = 0;
    for (size_t t = 0; t < terms.size(); ++t) {
        const auto& cls = terms[t];
        int s = 1, e = 1;
        if (!cls.empty()) {
            s = cls.front().start;
            e = cls.back().start + cls.back().width - 1;
        }
        term_end[t] = e;
        auto rk = std::make_pair(-s - 1, std::string());
        int cur;
        if (!index.count(rk)) {
            Node r;
            r.start = s;
            cur = (int)nodes.size();
            index[rk] = cur;
            nodes.push_back(r);
        } else {
            cur = index[rk];
        }
        size_t ci = 0;
        int i = s;
        while (i <= e) {
            const bool has = ci < cls.size() && cls[ci].start == i;
            const int w = has ? cls[ci].width : 1;
            std::string key = std::to_string(i) + ":" + std::to_string(w);
            if (has) key.append(reinterpret_cast<const char*>(cls[ci].op.data()), cls[ci].op.size() * sizeof(zc));
            auto k2 = std::make_pair(cur, key);
            auto it = index.find(k2);
            if (it == index.end()) {
                Node nd;
                nd.parent = cur;
                nd.depth = nodes[cur].depth + 1;
                nd.site = i;
                nd.width = w;
                nd.cl = has ? &cls[ci] : nullptr;
                nodes[cur].children += 1;
                const int id = (int)nodes.size();
                index[k2] = id;
                nodes.push_back(nd);
                cur = id;
                maxdepth = std::max(maxdepth, nd.depth);
            } else {
                cur = it->second;
            }
            i += w;
            if (has) ++ci;
        }
        nodes[cur].leaf = true;
        term_leaf[t] = cur;
    }
    std::vector<cplx*> L, R;
    lr_envs(c, m, L, R);
    for (auto& nd : nodes)
        if (nd.parent < 0) nd.E = L[nd.start - 1];
    std::vector<std::vector<int>> by_depth(maxdepth + 1);
    for (int k = 0; k < (int)nodes.size(); ++k) by_depth[nodes[k].depth].push_back(k);
    // operators of the clusters in device memory (one upload per distinct node)
    std::vector<const cplx*> dop(nodes.size(), nullptr);
    for (int k = 0; k < (int)nodes.size(); ++k)
        if (nodes[k].cl) {
            std::vector<cplx> h(nodes[k].cl->op.size());
            std::memcpy(h.data(), nodes[k].cl->op.data(), h.size() * sizeof(cplx));
            dop[k] = upload(c, h);
        }
    for (int depth = 1; depth <= maxdepth; ++depth) {
        const auto& lvl = by_depth[depth];
        // per node: the current buffer and its micro-operation counter
        std::vector<cplx*> cur(lvl.size());
        std::vector<int> curl(lvl.size());
        for (size_t q = 0; q < lvl.size(); ++q) {
            cur[q] = nodes[nodes[lvl[q]].parent].E;
            curl[q] = m->chi[nodes[lvl[q]].site - 1];
        }
        int maxw = 1;
        for (int k : lvl) maxw = std::max(maxw, nodes[k].width);
        const int phases = 2 * maxw + 1;  // w ket absorptions, the operator, w bra peels
        for (int ph = 0; ph < phases; ++ph) {
            std::vector<GemmItem> gm;
            std::vector<cplx*> to_free;
            std::vector<ApplyItem> ap;
            std::vector<int2> ap_tiles;
            for (size_t q = 0; q < lvl.size(); ++q) {
                const Node& nd = nodes[lvl[q]];
                const int i = nd.site, w = nd.width;
                const int lb = m->chi[i - 1], ra = m->chi[i + w - 1];
                // this node's phase ph: ket absorption k < w, then the operator, then bra peels
                const int nph = 2 * w + 1;
                if (ph >= nph) continue;
                if (ph < w) {
                    const int j = i + ph;
                    const int rows = lb * ipow(d, ph);
                    cplx* T = dalloc<cplx>(c, (size_t)rows * d * m->chi[j]);
                    GemmItem g = gemm_item(cur[q], GEMM_N, rows, m->site[j - 1], GEMM_N, m->chi[j - 1], T, rows, rows,
                                           d * m->chi[j], m->chi[j - 1]);
                    if (j >= 2) { g.dk = m->lam[j - 2]; g.dk_mod = m->chi[j - 1]; }
                    gm.push_back(g);
                    if (ph > 0) to_free.push_back(cur[q]);
                    cur[q] = T;
                } else if (ph == w) {
                    if (nd.cl) {
                        const size_t pairs = (size_t)lb * ra;
                        for (size_t b = 0; b * 256 < pairs; ++b) ap_tiles.push_back(make_int2((int)ap.size(), (int)b));
                        ap.push_back(ApplyItem{cur[q], lb, ipow(d, w), ra, dop[lvl[q]]});
                    }
                } else {
                    const int s2 = ph - w - 1;
                    const int j = i + s2;
                    const int rest = ipow(d, w - 1 - s2) * ra;
                    const int rb = m->chi[j];
                    cplx* W = dalloc<cplx>(c, (size_t)rb * rest);
                    GemmItem g = gemm_item(m->site[j - 1], GEMM_C, curl[q] * d, cur[q], GEMM_N, curl[q] * d, W, rb,
                                           rb, rest, curl[q] * d);
                    if (j >= 2) { g.dk = m->lam[j - 2]; g.dk_mod = curl[q]; }
                    gm.push_back(g);
                    to_free.push_back(cur[q]);
                    cur[q] = W;
                    curl[q] = rb;
                }
            }
            if (!ap.empty()) {
                ApplyItem* dap;
                int2* dapt;
                upload2(c, ap, ap_tiles, dap, dapt);
                launch_apply_op_batch(dap, dapt, (int)ap_tiles.size(), c->stream);
            }
            run_gemm(c, gm);
            for (cplx* p : to_free) dfree(c, p);
        }
        for (size_t q = 0; q < lvl.size(); ++q) nodes[lvl[q]].E = cur[q];
        // environments of the previous depth are no longer read, unless a term ends there
        for (int k : by_depth[depth - 1])
            if (nodes[k].parent >= 0 && !nodes[k].leaf) {
                dfree(c, nodes[k].E);
                nodes[k].E = nullptr;
            }
    }
    std::vector<CloseItem> close(terms.size());
    for (size_t t = 0; t < terms.size(); ++t) {
        const int e = term_end[t];
        const double* lb = e <= n - 1 ? m->lam[e - 1] : nullptr;
        close[t] = CloseItem{nodes[term_leaf[t]].E, R[e], m->chi[e], m->chi[e], lb, lb, vals + t};
    }
    launch_env_close_batch(upload(c, close), (int)close.size(), c->stream);
    for (auto& nd : nodes)
        if (nd.parent >= 0 && nd.E) dfree(c, nd.E);
    for (auto p : L) dfree(c, p);
    for (auto p : R) dfree(c, p);
}

}  // namespace

namespace {
// `Chain(::State, boundary, arrays)` (src/Ansatz/Chain.jl:36-62 periodic, :64-100 open)
int mps_create_impl(qh_context* c, int n, int d, const int* bond_dims, const double* const* arrays,
                    const double* const* lambdas, bool periodic, qh_mps** out) {
    if (!c || !out) return QH_ERR_ARGUMENT;
    qh_mps* m = nullptr;
    int rc = guard(c, [&] {
        if (n < 2) throw Error(QH_ERR_ARGUMENT, "a Chain needs at least 2 sites");
        if (d < 2 || d > 4) throw Error(QH_ERR_ARGUMENT, "physical dimension must be 2, 3 or 4");
        if (!bond_dims || !arrays) throw Error(QH_ERR_ARGUMENT, "null argument");
        const int nb = periodic ? n : n - 1;
        for (int b = 0; b < nb; ++b)
            if (bond_dims[b] < 1) throw Error(QH_ERR_DIMENSION, "bond dimensions must be positive");
        m = new qh_mps();
        m->ctx = c;
        m->n = n;
        m->d = d;
        m->periodic = periodic;
        m->chi.assign(n + 1, 1);
        for (int b = 1; b <= n - 1; ++b) m->chi[b] = bond_dims[b - 1];
        if (periodic) m->chi[0] = m->chi[n] = bond_dims[n - 1];
        m->site.assign(n, nullptr);
        m->lam.assign(n - 1, nullptr);
        std::vector<std::vector<cplx>> stage(n);
        for (int i = 1; i <= n; ++i) {
            if (!arrays[i - 1]) throw Error(QH_ERR_ARGUMENT, "null site array");
            const int cl = m->chi[i - 1], cr = m->chi[i];
            const size_t ne = site_elems(m, i);
            const cplx* src = reinterpret_cast<const cplx*>(arrays[i - 1]);
            auto& st = stage[i - 1];
            st.resize(ne);
            // (o, l, r) -> (l, o, r)
            for (int r = 0; r < cr; ++r)
                for (int l = 0; l < cl; ++l)
                    for (int o = 0; o < d; ++o)
                        st[l + (size_t)cl * (o + (size_t)d * r)] = src[o + (size_t)d * (l + (size_t)cl * r)];
            m->site[i - 1] = dalloc<cplx>(c, ne);
            keep(c, m->site[i - 1]);
            QH_HIP(hipMemcpyAsync(m->site[i - 1], st.data(), ne * sizeof(cplx), hipMemcpyHostToDevice, c->stream));
        }
        if (lambdas) {
            for (int b = 1; b <= n - 1; ++b) {
                if (!lambdas[b - 1]) continue;
                m->lam[b - 1] = dalloc<double>(c, m->chi[b]);
                keep(c, m->lam[b - 1]);
                QH_HIP(hipMemcpyAsync(m->lam[b - 1], lambdas[b - 1], m->chi[b] * sizeof(double), hipMemcpyHostToDevice,
                                      c->stream));
            }
        }
        ctx_sync(c);
    });
    if (rc != QH_OK) {
        if (m) {
            for (auto p : m->site) dfree(c, p);
            for (auto p : m->lam) dfree(c, p);
            delete m;
        }
        return rc;
    }
    *out = m;
    return QH_OK;
}
}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int qh_ctx_create(int device, qh_context** out) {
    if (!out) { thread_error() = "null output pointer"; return QH_ERR_ARGUMENT; }
    qh_context* c = nullptr;
    int rc = guard(nullptr, [&] {
        int ndev = 0;
        QH_HIP(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw Error(QH_ERR_ARGUMENT, "invalid device ordinal");
        QH_HIP(hipSetDevice(device));
        noalloc_setup(device);  // this device's scratch address for qh_overlap_noalloc
        c = new qh_context();
        c->device = device;
        {
            int ncu = 0;
            if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
                c->num_cus = ncu;
        }
        QH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        if (const char* e = std::getenv("QH_GRAPHS")) c->use_graphs = std::atoi(e) != 0;
        if (const char* e = std::getenv("QH_PIVOT")) c->pivot = std::atoi(e) != 0;
        if (const char* e = std::getenv("QH_JACOBI_TAIL")) c->jac_tail = std::atoi(e) != 0;
        if (const char* e = std::getenv("QH_JACOBI_DEFER_V")) c->jac_defer_v = std::atoi(e) != 0;
        if (const char* e = std::getenv("QH_JACOBI_TAIL_SWITCH"))
            c->jac_tail_switch = std::min(1000, std::max(1, std::atoi(e)));
        if (const char* e = std::getenv("QH_JACOBI_BLOCK")) {
            const int b = std::atoi(e);
            if (b == 8 || b == 16 || b == 32) c->jac_block = b;
        }
        if (const char* e = std::getenv("QH_QR_PASSES")) c->jac_precondition = std::min(std::max(std::atoi(e), 0), QR_MAX_PASSES);
        // a small first chunk makes tests exercise the multi-chunk arena cheaply
        if (const char* e = std::getenv("QH_ARENA_INITIAL_BYTES")) {
            const long long v = std::atoll(e);
            if (v >= 256) c->arena_first = (size_t)v;
        }
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
            uint64_t thr = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
        }
    });
    if (rc != QH_OK) {
        if (c) {
            if (c->stream) (void)hipStreamDestroy(c->stream);
            delete c;
        }
        return rc;
    }
    *out = c;
    return QH_OK;
}

int qh_ctx_destroy(qh_context* c) {
    if (!c) return QH_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& p : c->pending) { c->event_pool.push_back(p.a); c->event_pool.push_back(p.b); }
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    arena_free(c);
    pool_destroy(c);
    if (c->h_read) (void)hipHostFree(c->h_read);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return QH_OK;
}

int qh_ctx_synchronize(qh_context* c) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] { ctx_sync(c); });
}

const char* qh_last_error(const qh_context* c) { return c ? c->err.c_str() : thread_error().c_str(); }

int qh_ctx_set_timing(qh_context* c, int enable) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        ctx_sync(c);
        c->timing = enable != 0;
    });
}

int qh_ctx_get_timing(qh_context* c, const char* cls, double* ms, int64_t* launches, double* flops, double* bytes) {
    if (!c || !cls) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        ctx_sync(c);
        auto it = c->timers.find(cls);
        TimerClass t = it == c->timers.end() ? TimerClass{} : it->second;
        if (ms) *ms = t.ms;
        if (launches) *launches = t.launches;
        if (flops) *flops = t.flops;
        if (bytes) *bytes = t.bytes;
    });
}

int qh_ctx_reset_timing(qh_context* c) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        ctx_sync(c);
        c->timers.clear();
    });
}

int qh_ctx_set_jacobi(qh_context* c, int block, int max_sweeps) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        if (block != 0 && block != 8 && block != 16 && block != 32)
            throw Error(QH_ERR_ARGUMENT, "Jacobi block must be 8, 16 or 32");
        if (block) {
            c->jac_block = block;
            c->tuned_block.clear();  // an explicit setting replaces the tuned one
        }
        if (max_sweeps > 0) c->jac_max_sweeps = max_sweeps;
        if (const char* e = std::getenv("QH_JACOBI_MAX_INNER")) c->jac_max_inner = std::max(1, std::atoi(e));
    });
}

int qh_ctx_set_qr_passes(qh_context* c, int passes) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        if (passes < 0 || passes > QR_MAX_PASSES) throw Error(QH_ERR_ARGUMENT, "QR passes must be in 0..4");
        c->jac_precondition = passes;
        c->tuned_passes.clear();  // an explicit setting replaces the tuned one
    });
}

// The SVD schedule tuner.  For each size class q = 64, 128, ... <= max_q it factors a
// batch of q × q synthetic matrices -- half with i.i.d. uniform entries (a flat
// spectrum, like Haar-random gates on full-rank states), half products A·D·B of
// uniform matrices with D decaying from 1 to 1e-12 (a graded spectrum whose singular
// vectors are not aligned with the columns, like a truncated time evolution)
// -- through the batched SVD under each candidate schedule (Jacobi block width 8 or
// 16, 1..4 QR preconditioning passes; width 32 is left out: a batch runs at its
// largest class's width, and width 32 slows the small matrices of mixed batches), timed end to end after one warm-up run, and
// keeps the fastest for the class.  Batches shrink with q² above q = 512 so every
// class costs about the same.  The candidates change the rounding of the results,
// never their meaning.
int qh_ctx_autotune(qh_context* c, int max_q, int batch, double* report, int* nclasses) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        if (max_q < 64 || max_q > 4096) throw Error(QH_ERR_ARGUMENT, "max_q must be in 64..4096");
        if (batch < 2 || batch > 512) throw Error(QH_ERR_ARGUMENT, "batch must be in 2..512");
        const int keep_block = c->jac_block, keep_passes = c->jac_precondition;
        std::map<int, int> passes_tab, block_tab;
        c->tuned_passes.clear();
        c->tuned_block.clear();
        int ncls = 0;
        for (int q = 64; q <= max_q; q *= 2, ++ncls) {
            const int b = std::max(2, q <= 512 ? batch : (int)((long long)batch * 512 * 512 / ((long long)q * q)));
            const size_t elems = (size_t)q * q;
            std::vector<cplx*> src(b), work(b), factors;
            std::vector<double*> lam(b);
            std::vector<GemmItem> prod;
            for (int i = 0; i < b; ++i) {
                src[i] = dalloc<cplx>(c, elems);
                work[i] = dalloc<cplx>(c, elems);
                lam[i] = dalloc<double>(c, q);
                const unsigned long long seed = 0x5eedull + (unsigned long long)q;
                if (i % 2 == 0) {
                    launch_rand_uniform(src[i], elems, seed, (unsigned long long)i, 1, c->stream);
                } else {  // A diag(10^(-12 j/(q-1))) B: a graded spectrum in no particular basis
                    cplx* fa = dalloc<cplx>(c, elems);
                    cplx* fb = dalloc<cplx>(c, elems);
                    factors.push_back(fa);
                    factors.push_back(fb);
                    launch_rand_uniform(fa, elems, seed, 2ull * i, 1, c->stream);
                    launch_rand_uniform(fb, elems, seed, 2ull * i + 1, 1, c->stream);
                    launch_scale_cols_geometric(fa, q, q, 12.0, c->stream);
                    prod.push_back(gemm_item(fa, GEMM_N, q, fb, GEMM_N, q, src[i], q, q, q, q));
                }
            }
            run_gemm(c, prod, QH_PREC_FP64);
            for (cplx* f : factors) dfree(c, f);
            auto run = [&](int block, int passes) {
                c->jac_block = block;
                c->jac_precondition = passes;
                for (int i = 0; i < b; ++i)
                    QH_HIP(hipMemcpyAsync(work[i], src[i], elems * sizeof(cplx), hipMemcpyDeviceToDevice, c->stream));
                std::vector<SvdTask> tasks(b);
                for (int i = 0; i < b; ++i) {
                    tasks[i].x = work[i];
                    tasks[i].mr = tasks[i].mc = q;
                    tasks[i].truncate = 0;
                    tasks[i].lam_dst = lam[i];
                }
                ctx_sync(c);
                const auto t0 = std::chrono::steady_clock::now();
                run_svd_tasks(c, tasks);  // synchronises
                return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            };
            (void)run(16, 3);  // warm-up: allocations of the pool, first launches
            // the schedule the context would use without the tuner
            const int dflt_block = keep_block > 0 ? keep_block : (q >= 2048 ? 32 : 16);
            const double dflt = run(dflt_block, keep_passes);
            double best = dflt;
            int bb = dflt_block, bp = keep_passes;
            for (int block : {8, 16})
                for (int passes = 1; passes <= QR_MAX_PASSES; ++passes) {
                    const double ms = run(block, passes);
                    if (ms < best) { best = ms; bb = block; bp = passes; }
                }
            // a candidate within 3 % of the current schedule is timing noise on synthetic
            // matrices, not a gain on real ones: the current schedule stays
            if (best > 0.97 * dflt) {
                best = dflt;
                bb = dflt_block;
                bp = keep_passes;
            }
            passes_tab[q] = bp;
            block_tab[q] = bb;
            if (report && ncls < 8) {
                double* r = report + 5 * ncls;
                r[0] = q; r[1] = bb; r[2] = bp; r[3] = best; r[4] = dflt;
            }
            for (int i = 0; i < b; ++i) {
                dfree(c, src[i]);
                dfree(c, work[i]);
                dfree(c, lam[i]);
            }
        }
        c->jac_block = keep_block;
        c->jac_precondition = keep_passes;
        c->tuned_passes = passes_tab;
        c->tuned_block = block_tab;
        if (nclasses) *nclasses = std::min(ncls, 8);
    });
}

int qh_ctx_set_graphs(qh_context* c, int enable) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] { c->use_graphs = enable != 0; });
}

int qh_ctx_set_pivoting(qh_context* c, int enable) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] { c->pivot = enable != 0; });
}

// forget the tuned schedule (back to qh_ctx_set_jacobi / qh_ctx_set_qr_passes)
int qh_ctx_autotune_reset(qh_context* c) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        c->tuned_passes.clear();
        c->tuned_block.clear();
    });
}

int qh_ctx_mem_info(qh_context* c, size_t* free_bytes, size_t* total_bytes) {
    if (!c || !free_bytes || !total_bytes) return QH_ERR_ARGUMENT;
    return guard(c, [&] { QH_HIP(hipMemGetInfo(free_bytes, total_bytes)); });
}

int qh_mps_create(qh_context* c, int n, int d, const int* bond_dims, const double* const* arrays,
                  const double* const* lambdas, qh_mps** out) {
    return mps_create_impl(c, n, d, bond_dims, arrays, lambdas, false, out);
}

int qh_pmps_create(qh_context* c, int n, int d, const int* bond_dims, const double* const* arrays, qh_mps** out) {
    return mps_create_impl(c, n, d, bond_dims, arrays, nullptr, true, out);
}

// `rand(Chain, Open, State; n, χ, p, eltype)` (src/Ansatz/Chain.jl:223-256): bond b
// has dimension min(χ, p^b, p^(n-b)); every site but the first is a right isometry
// (orthonormal rows of a matrix of uniform [0, 1) entries, `gramschmidt!` at
// Chain.jl:242), the first is p × χ1 with orthonormal columns divided by sqrt(p), so
// the state has norm 1 when χ >= p.  The orthonormalisation is the batched
// Householder QR of the transposed random matrices (its phases differ from
// Gram-Schmidt's, a gauge), the entries a counter-based hash of (seed, site, entry).
int qh_mps_rand(qh_context* c, int n, int chi, int p, int complex_entries, unsigned long long seed, qh_mps** out) {
    if (!c || !out) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        if (n < 2) throw Error(QH_ERR_ARGUMENT, "a Chain needs at least 2 sites");
        if (p < 2 || p > 4) throw Error(QH_ERR_ARGUMENT, "physical dimension must be 2, 3 or 4");
        if (chi < 1) throw Error(QH_ERR_ARGUMENT, "χ must be positive");
        std::unique_ptr<qh_mps> m(new qh_mps());
        m->ctx = c;
        m->n = n;
        m->d = p;
        m->chi.assign(n + 1, 1);
        for (int b = 1; b <= n - 1; ++b) {
            long long v = 1;
            for (int e = 0; e < std::min(b, n - b) && v < chi; ++e) v *= p;
            m->chi[b] = (int)std::min<long long>(v, chi);
        }
        m->site.assign(n, nullptr);
        m->lam.assign(n - 1, nullptr);
        // site 1: X (p × χ1); site i >= 2: X = S^H ((p χ_i) × χ_{i-1}), S the site as χl × (p χr)
        std::vector<QrJob> jobs(n);
        std::vector<cplx*> q(n);
        for (int i = 1; i <= n; ++i) {
            const int rows = (i == 1) ? p : p * m->chi[i];
            const int cols = (i == 1) ? m->chi[1] : m->chi[i - 1];
            cplx* x = dalloc<cplx>(c, (size_t)rows * cols);
            launch_rand_uniform(x, (size_t)rows * cols, seed, (unsigned long long)i, complex_entries, c->stream);
            const size_t tq = (size_t)((cols + QR_NB - 1) / QR_NB) * QR_NB * QR_NB;
            jobs[i - 1] = QrJob{x, rows, cols, rows, dalloc<cplx>(c, (size_t)rows * cols), dalloc<cplx>(c, tq)};
            q[i - 1] = dalloc<cplx>(c, (size_t)rows * cols);
            launch_eye_rect(q[i - 1], rows, cols, c->stream);
        }
        run_qr_factor(c, jobs);
        std::vector<QrApplyJob> app;
        for (int i = 1; i <= n; ++i) app.push_back(QrApplyJob{&jobs[i - 1], q[i - 1], jobs[i - 1].n});
        run_qr_apply_q(c, app);
        for (int i = 1; i <= n; ++i) {
            const QrJob& j = jobs[i - 1];
            if (i == 1) {
                launch_scale_real(q[0], (size_t)j.m * j.n, 1.0 / std::sqrt((double)p), c->stream);
                m->site[0] = q[0];  // (1, o, r) = Q[o, r] / sqrt(p)
            } else {
                m->site[i - 1] = dalloc<cplx>(c, (size_t)j.m * j.n);
                launch_conj_transpose(q[i - 1], j.m, j.n, m->site[i - 1], c->stream);  // S = Q^H
                dfree(c, q[i - 1]);
            }
            dfree(c, j.a);
            dfree(c, j.v);
            dfree(c, j.t);
        }
        for (auto s2 : m->site) keep(c, s2);
        *out = m.release();
    });
}

int qh_mps_is_periodic(const qh_mps* m, int* periodic, int* wrap_dim) {
    if (!m || !periodic || !wrap_dim) return QH_ERR_ARGUMENT;
    *periodic = m->periodic ? 1 : 0;
    *wrap_dim = m->chi[0];
    return QH_OK;
}

int qh_mps_clone(qh_context* c, const qh_mps* src, qh_mps** out) {
    if (!c || !out) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        check_mps(c, src);
        // buffers stay call-owned until the copy is complete: a failure frees them all
        std::unique_ptr<qh_mps> m(new qh_mps(*src));
        for (int i = 1; i <= m->n; ++i) m->site[i - 1] = dcopy(c, src->site[i - 1], site_elems(src, i));
        for (int b = 1; b <= m->n - 1; ++b)
            m->lam[b - 1] = src->lam[b - 1] ? dcopy_d(c, src->lam[b - 1], src->chi[b]) : nullptr;
        for (auto p : m->site) keep(c, p);
        for (auto p : m->lam) keep(c, p);
        *out = m.release();
    });
}

int qh_mps_free(qh_mps* m) {
    if (!m) return QH_OK;
    qh_context* c = m->ctx;
    int rc = guard(c, [&] {
        for (auto p : m->site) dfree(c, p);
        for (auto p : m->lam) dfree(c, p);
    });
    delete m;
    return rc;
}

int qh_mps_nsites(const qh_mps* m, int* n, int* d) {
    if (!m) return QH_ERR_ARGUMENT;
    if (n) *n = m->n;
    if (d) *d = m->d;
    return QH_OK;
}

int qh_mps_bond_dims(qh_mps* m, int* dims, int* has) {
    if (!m) return QH_ERR_ARGUMENT;
    for (int b = 1; b <= m->n - 1; ++b) {
        if (dims) dims[b - 1] = m->chi[b];
        if (has) has[b - 1] = m->lam[b - 1] ? 1 : 0;
    }
    return QH_OK;
}

int qh_mps_get_site(qh_mps* m, int site, double* host) {
    if (!m || !host) return QH_ERR_ARGUMENT;
    qh_context* c = m->ctx;
    return guard(c, [&] {
        check_site(m, site);
        const int cl = m->chi[site - 1], cr = m->chi[site], d = m->d;
        const size_t ne = site_elems(m, site);
        std::vector<cplx> tmp(ne);
        readback(c, tmp.data(), m->site[site - 1], ne * sizeof(cplx));
        cplx* dst = reinterpret_cast<cplx*>(host);
        for (int r = 0; r < cr; ++r)
            for (int l = 0; l < cl; ++l)
                for (int o = 0; o < d; ++o)
                    dst[o + (size_t)d * (l + (size_t)cl * r)] = tmp[l + (size_t)cl * (o + (size_t)d * r)];
    });
}

int qh_mps_get_lambda(qh_mps* m, int bond, double* host) {
    if (!m || !host) return QH_ERR_ARGUMENT;
    qh_context* c = m->ctx;
    return guard(c, [&] {
        check_bond(m, bond);
        if (!m->lam[bond - 1])
            throw Error(QH_ERR_MISSING_SCHMIDT, "Can't access the spectrum on bond (" + std::to_string(bond) + ", " +
                                                    std::to_string(bond + 1) + ")");
        readback(c, host, m->lam[bond - 1], m->chi[bond] * sizeof(double));
    });
}

int qh_mps_footprint(const qh_mps* m, size_t* bytes) {
    if (!m || !bytes) return QH_ERR_ARGUMENT;
    size_t b = 0;
    for (int i = 1; i <= m->n; ++i) b += site_elems(m, i) * sizeof(cplx);
    for (int k = 1; k <= m->n - 1; ++k)
        if (m->lam[k - 1]) b += m->chi[k] * sizeof(double);
    *bytes = b;
    return QH_OK;
}

int qh_evolve1(qh_context* c, qh_mps* m, int site, const double* gate) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        check_mps(c, m);
        if (!gate) throw Error(QH_ERR_ARGUMENT, "null gate");
        if (site < 1 || site > m->n) throw Error(QH_ERR_ARGUMENT, "Gate inputs must be a subset of the TN sites");
        const int d = m->d;
        std::vector<cplx> hg((size_t)d * d);
        std::memcpy(hg.data(), gate, hg.size() * sizeof(cplx));
        const cplx* dg = upload(c, hg);
        Timed tm(c, "apply1");
        launch_apply_op(m->site[site - 1], m->chi[site - 1], d, m->chi[site], dg, c->stream);
    });
}

int qh_evolve2(qh_context* c, qh_mps* m, int bond, const double* gate, double threshold, int maxdim, unsigned flags,
               int* kept, double* discarded) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] { evolve2_core(c, 1, &m, &bond, gate, &threshold, &maxdim, flags, kept, discarded); });
}

int qh_evolve2_batch(qh_context* c, int count, qh_mps* const* ms, const int* bonds, const double* gates,
                     const double* thresholds, const int* maxdims, unsigned flags, int* kept, double* discarded) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] { evolve2_core(c, count, ms, bonds, gates, thresholds, maxdims, flags, kept, discarded); });
}

int qh_truncate(qh_context* c, qh_mps* m, int bond, double threshold, int maxdim) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        check_mps(c, m);
        truncate_core(c, m, bond, threshold, maxdim);
    });
}

int qh_contract_between(qh_context* c, qh_mps* m, int s1, int s2, int direction, int del) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        check_mps(c, m);
        contract_between_core(c, m, s1, s2, direction, del);
    });
}

int qh_canonize_site(qh_context* c, qh_mps* m, int site, int direction, int method) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        check_mps(c, m);
        canonize_site_core(c, m, site, direction, method);
    });
}

int qh_canonize(qh_context* c, qh_mps* m) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        check_mps(c, m);
        canonize_core(c, m);
    });
}

int qh_mixed_canonize(qh_context* c, qh_mps* m, int center) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        check_mps(c, m);
        mixed_canonize_core(c, m, center);
    });
}

int qh_normalize(qh_context* c, qh_mps* m, int root) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        check_mps(c, m);
        check_site(m, root);
        normalize_core(c, m, root);
    });
}

static void canon_test(qh_context* c, qh_mps* m, int site, double atol, int* result, double* dev, bool left) {
    check_mps(c, m);
    check_site(m, site);
    const int cl = m->chi[site - 1], cr = m->chi[site], d = m->d;
    const int nn = left ? cr : cl;
    cplx* g = dalloc<cplx>(c, (size_t)nn * nn);
    std::vector<GemmItem> gi;
    if (left) gi.push_back(gemm_item(m->site[site - 1], 1, cl * d, m->site[site - 1], 0, cl * d, g, cr, cr, cr, cl * d));
    else gi.push_back(gemm_item(m->site[site - 1], 0, cl, m->site[site - 1], 1, cl, g, cl, cl, cl, d * cr));
    run_gemm(c, gi);
    double* dd = dalloc<double>(c, 1);
    launch_gram_dev(g, nn, dd, c->stream);
    double h = 0.0;
    readback(c, &h, dd, sizeof(double));
    dfree(c, g);
    dfree(c, dd);
    if (result) *result = (h <= atol) ? 1 : 0;
    if (dev) *dev = h;
}

int qh_is_left_canonical(qh_context* c, qh_mps* m, int site, double atol, int* result, double* dev) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] { canon_test(c, m, site, atol, result, dev, true); });
}

int qh_is_right_canonical(qh_context* c, qh_mps* m, int site, double atol, int* result, double* dev) {
    if (!c) return QH_ERR_ARGUMENT;
    return guard(c, [&] { canon_test(c, m, site, atol, result, dev, false); });
}

int qh_expect(qh_context* c, qh_mps* m, int nobs, const qh_observable* obs, double* out) {
    if (!c || !out) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        check_mps(c, m);
        if (nobs < 0 || (nobs > 0 && !obs)) throw Error(QH_ERR_ARGUMENT, "bad observable list");
        auto cls = make_clusters(m, nobs, obs);
        zc v = full_contraction(c, m, m, cls);
        out[0] = v.real();
        out[1] = v.imag();
    });
}

int qh_expect_many(qh_context* c, qh_mps* m, int nterms, const int* offsets, const qh_observable* obs, double* out) {
    if (!c || !out || !offsets) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        check_mps(c, m);
        std::vector<std::vector<Cluster>> terms(nterms);
        for (int t = 0; t < nterms; ++t) {
            if (offsets[t + 1] < offsets[t]) throw Error(QH_ERR_ARGUMENT, "bad term offsets");
            terms[t] = make_clusters(m, offsets[t + 1] - offsets[t], obs + offsets[t]);
        }
        if (m->periodic) {  // no shared environments: one periodic contraction per term
            for (int t = 0; t < nterms; ++t) {
                const zc v = periodic_contraction(c, m, m, terms[t]);
                out[2 * t] = v.real();
                out[2 * t + 1] = v.imag();
            }
            return;
        }
        cplx* vals = dalloc<cplx>(c, std::max(nterms, 1));
        expect_terms_batched(c, m, terms, vals);
        std::vector<cplx> hv(std::max(nterms, 1));
        readback(c, hv.data(), vals, hv.size() * sizeof(cplx));
        std::memcpy(out, hv.data(), (size_t)nterms * sizeof(cplx));
        dfree(c, vals);
    });
}

int qh_overlap(qh_context* c, qh_mps* a, qh_mps* b, double* out) {
    if (!c || !out) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        check_mps(c, a);
        check_mps(c, b);
        if (a->n != b->n || a->d != b->d) throw Error(QH_ERR_ASSERTION, "Ansatzes must have the same sites");
        zc v = full_contraction(c, a, b, {});
        out[0] = v.real();
        out[1] = v.imag();
    });
}

int qh_norm(qh_context* c, qh_mps* m, double* out) {
    if (!c || !out) return QH_ERR_ARGUMENT;
    return guard(c, [&] {
        check_mps(c, m);
        zc v = full_contraction(c, m, m, {});
        *out = std::abs(std::sqrt(v));
    });
}

}  // extern "C"
