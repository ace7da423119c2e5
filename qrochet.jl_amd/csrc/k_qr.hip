// Batched blocked Householder QR (complex FP64), used to precondition the
// one-sided Jacobi SVD:  X = Q1 R1,  R_{j-1}^H = Q_j R_j,  Jacobi on R_n^H (three
// passes by default).  The triangular factors grade the columns so that the
// Jacobi sweeps converge in fewer sweeps (the preconditioning of Drmac & Veselic,
// "New fast and accurate Jacobi SVD algorithm", SIAM J. Matrix Anal. Appl. 29
// (2008); on the matrices of a TEBD layer, 512 x 512, two passes take the rounds
// from ~1600 to ~900 per three layers and a third pass to ~775).
//
// Panels of QR_NB columns: one workgroup factors a panel (LAPACK zlarfg
// reflectors, H^H x = beta e1 with H = I - tau v v^H, real beta) and builds the
// compact-WY triangle T (Q_panel = I - V T V^H); a second kernel applies
// Q_panel^H (factorisation) or Q_panel (forming Q·M) to column tiles.
#include "qh_kernels.h"
#include "qh_complex.h"

namespace qh {

constexpr int QR_TC = 16;  // target columns per apply workgroup

__device__ __forceinline__ double2 shfl_xor2(double2 a, int off) {
    return make_double2(__shfl_xor(a.x, off, 64), __shfl_xor(a.y, off, 64));
}

// the value of lane `src` (the same lane for the whole wave) in every lane
__device__ __forceinline__ double bcast_d(double x, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), src);
    return __hiloint2double(hi, lo);
}

// Cross-lane swaps of gfx950 (per dword of a double): permlane32 trades lanes 32..63
// of a with lanes 0..31 of b, permlane16 trades the odd 16-lane rows of a (lanes
// 16..31, 48..63) with the even rows of b (lanes 0..15, 32..47).  Afterwards a + b
// holds, in every lane, its own value plus the partner's (lane ^ 32, resp. ^ 16) of the
// half that the lane keeps: a's for the lower lanes, b's for the upper ones.
__device__ __forceinline__ void swap_lanes32(double& a, double& b) {
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(a), (unsigned)__double2loint(b),
                                                     false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(a), (unsigned)__double2hiint(b),
                                                     false, false);
    a = __hiloint2double((int)hi[0], (int)lo[0]);
    b = __hiloint2double((int)hi[1], (int)lo[1]);
}
__device__ __forceinline__ void swap_lanes16(double& a, double& b) {
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(a), (unsigned)__double2loint(b),
                                                     false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(a), (unsigned)__double2hiint(b),
                                                     false, false);
    a = __hiloint2double((int)hi[0], (int)lo[0]);
    b = __hiloint2double((int)hi[1], (int)lo[1]);
}
// halving step of block_sum16 with a lane swap: slot i (lower lanes) or i + half
// (upper lanes), summed over the lane and its partner, lands in v[i]
template <int OFF>
__device__ __forceinline__ double2 halve_swap(double2 lo, double2 hi) {
    if constexpr (OFF == 32) {
        swap_lanes32(lo.x, hi.x);
        swap_lanes32(lo.y, hi.y);
    } else {
        swap_lanes16(lo.x, hi.x);
        swap_lanes16(lo.y, hi.y);
    }
    return cadd(lo, hi);
}

// LDS written by one wave and read back by the same wave: the fences order the
// accesses across the wave's lanes, and no workgroup barrier is needed
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum over the workgroup (NW waves) of 16 complex values per thread, with one barrier.
// Each wave halves its values recursively: at the step with offset o a lane keeps the
// upper or the lower half (by lane bit o) and adds the partner's copy of that half (lane
// swaps for o = 32, 16; shuffles for o = 8, 4), so 16 values take 15 exchanges; lane l
// then holds value (l >> 2) & 15 summed over the
// lanes that agree with l in bits 0-1, which two plain steps add.  The wave's 16 sums go
// to sRed; after the barrier every lane adds the NW partials of value l & 15 in wave
// order, so all waves hold the same totals, value k in lanes k, k + 16, k + 32, k + 48.
// sRed is read after the barrier, so consecutive calls alternate two buffers: a fast
// wave's next writes then cannot overtake a slow wave's reads.
template <int NW>
__device__ __forceinline__ double2 block_sum16(double2 (&v)[16], double2* sRed /* NW*16 */) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double2 h8[8], h4[4];
#pragma unroll
    for (int i = 0; i < 8; ++i) h8[i] = halve_swap<32>(v[i], v[i + 8]);
#pragma unroll
    for (int i = 0; i < 4; ++i) h4[i] = halve_swap<16>(h8[i], h8[i + 4]);
    double2 h2[2];
    {
        const bool up = (lane & 8) != 0;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const double2 send = up ? h4[i] : h4[i + 2];
            const double2 keep = up ? h4[i + 2] : h4[i];
            h2[i] = cadd(keep, shfl_xor2(send, 8));
        }
    }
    double2 s;
    {
        const bool up = (lane & 4) != 0;
        const double2 send = up ? h2[0] : h2[1];
        const double2 keep = up ? h2[1] : h2[0];
        s = cadd(keep, shfl_xor2(send, 4));
    }
    s = cadd(s, shfl_xor2(s, 2));
    s = cadd(s, shfl_xor2(s, 1));
    if ((lane & 3) == 0) sRed[wv * 16 + ((lane >> 2) & 15)] = s;
    __syncthreads();
    double2 t = sRed[lane & 15];
#pragma unroll
    for (int w = 1; w < NW; ++w) t = cadd(t, sRed[w * 16 + (lane & 15)]);
    return t;
}

// LAPACK zlarfg: H^H (alpha; x) = (beta; 0), H = I - tau v v^H, v = (1; x·scale), beta real
struct Reflector {
    double2 tau, scale;
    double beta;
};
__device__ __forceinline__ Reflector make_reflector(double2 alpha, double xnorm2) {
    Reflector h;
    h.tau = cz();
    h.scale = cz();
    h.beta = alpha.x;
    if (!(xnorm2 == 0.0 && alpha.y == 0.0)) {
        const double nrm = sqrt(cabs2(alpha) + xnorm2);
        h.beta = alpha.x >= 0.0 ? -nrm : nrm;
        const double rb = 1.0 / h.beta;
        h.tau = make_double2((h.beta - alpha.x) * rb, -alpha.y * rb);
        const double2 den = make_double2(alpha.x - h.beta, alpha.y);
        const double rd = 1.0 / cabs2(den);
        h.scale = make_double2(den.x * rd, -den.y * rd);  // 1 / den
    }
    return h;
}

// The column step of a panel (column c, rows relative to the panel corner): one
// workgroup sum gives, over the rows below the diagonal with x = column c there,
// s_c = ||x||^2, s_k = x^H a_k (k > c) and s_k = v_k^H x (k < c); row c (alpha = a_c(c),
// a_k(c), v_k(c)) is published by its owner before that sum's barrier.  With the
// reflector v = (1; x·scale) the dots follow without a second sum:
// v^H a_k = a_k(c) + conj(scale)·s_k and v_k^H v = conj(v_k(c)) + scale·s_k.
struct PanelShared {
    double2 sRed[2][8 * QR_NB];  // wave partial sums (up to 8 waves), two alternating buffers
    double2 sRow[2][QR_NB];      // row c of the panel, two alternating buffers
    double2 sBc[8][QR_NB];       // per wave: the 16 dots, written by lanes 0..15 and read by all
    double2 sT[QR_NB][QR_NB + 1];
};

// the coefficient of slot k (held by lane k of every wave) for every lane: lanes 0..15 write
// the wave's own LDS row, which the wave then reads where it uses a dot (the next
// column's writes follow these reads in the wave's program order, after a barrier)
__device__ __forceinline__ const double2* panel_dots_publish(double2 dl, PanelShared& sh) {
    double2* bc = sh.sBc[threadIdx.x >> 6];
    if ((threadIdx.x & 63) < QR_NB) bc[threadIdx.x & 63] = dl;
    wave_lds_sync();
    return bc;
}

// slot k's coefficient from the unscaled sum `tot` and row c's entry `row`: for k > c
// the trailing update's conj(tau) v^H a_k, for k < c the dot v_k^H v that builds T,
// 0 for k = c
__device__ __forceinline__ double2 panel_coef(int k, int c, double2 tot, double2 row, const Reflector& h) {
    if (k > c) return cmul(cconj(h.tau), cadd(row, cmul(cconj(h.scale), tot)));
    if (k < c) return cadd(cconj(row), cmul(h.scale, tot));
    return cz();
}

// Column loop of a panel factorisation on a panel view P(c, r) (column c, row r
// relative to the panel corner) in global memory, NT threads.  Leaves R on and
// above the diagonal, the scaled reflectors below it, and the compact-WY triangle
// in sh.sT.  Thread tid owns rows r = tid + NT j for the whole panel, so the sums,
// the reflector's scaling and the trailing update touch only the thread's own rows;
// the only cross-thread data are the one sum per column and row c (PanelShared).
template <int NT, class Ref>
__device__ void qr_panel_columns(int rows, int nb, PanelShared& sh, Ref P) {
    constexpr int NW = NT / 64;
    const int tid = threadIdx.x;
    for (int c = 0; c < nb; ++c) {
        double2 sk[QR_NB];
#pragma unroll
        for (int k = 0; k < QR_NB; ++k) sk[k] = cz();
        for (int r = tid; r < rows; r += NT) {
            if (r <= c) continue;
            const double2 x = P(c, r);
            // all 16 loads issue before their first use (column min(k, nb - 1) is always
            // inside the panel; the values of k >= nb are discarded)
            double2 pk[QR_NB];
#pragma unroll
            for (int k = 0; k < QR_NB; ++k) pk[k] = P(min(k, nb - 1), r);
#pragma unroll
            for (int k = 0; k < QR_NB; ++k) {
                if (k < c) sk[k] = cfmac(pk[k], x, sk[k]);
                else if (k == c) sk[k].x += cabs2(x);
                else if (k < nb) sk[k] = cfmac(x, pk[k], sk[k]);
            }
        }
        if (tid == c) {
#pragma unroll
            for (int k = 0; k < QR_NB; ++k) sh.sRow[c & 1][k] = (k < nb && c < rows) ? P(k, c) : cz();
        }
        const double2 tot = block_sum16<NW>(sk, sh.sRed[c & 1]);
        const Reflector h = make_reflector(sh.sRow[c & 1][c], bcast_d(tot.x, c));
        const double2 dl = panel_coef(tid & 15, c, tot, sh.sRow[c & 1][tid & 15], h);
        const double2* dk = panel_dots_publish(dl, sh);
        if (tid == c && c < rows) P(c, c) = make_double2(h.beta, 0.0);
        if (tid < QR_NB) {
            const int i = tid;
            double2 val = cz();
            if (i < c) {
                double2 acc = cz();
#pragma unroll
                for (int l = 0; l < QR_NB; ++l)
                    if (l >= i && l < c) acc = cfma(sh.sT[i][l], dk[l], acc);
                const double2 tc = h.tau;
                val = make_double2(-(tc.x * acc.x - tc.y * acc.y), -(tc.x * acc.y + tc.y * acc.x));
            } else if (i == c) {
                val = h.tau;
            }
            sh.sT[i][c] = val;
        }
        // v below the diagonal (scaled in place) and the trailing update
        // a_k -= conj(tau) (v^H a_k) v on own rows
        double2 w[QR_NB];
#pragma unroll
        for (int k = 0; k < QR_NB; ++k) w[k] = (k > c && k < nb) ? dk[k] : cz();
        for (int r = tid; r < rows; r += NT) {
            if (r < c) continue;
            double2 vr = make_double2(1.0, 0.0);
            if (r > c) {
                vr = cmul(P(c, r), h.scale);
                P(c, r) = vr;
            }
            double2 pk[QR_NB];
#pragma unroll
            for (int k = 0; k < QR_NB; ++k) pk[k] = P(min(k, nb - 1), r);
#pragma unroll
            for (int k = 0; k < QR_NB; ++k)
                if (k > c && k < nb) P(k, r) = cfms(w[k], vr, pk[k]);
        }
    }
    __syncthreads();
}

// panel factorisation in place in global memory (any number of rows; the panel's
// 16 columns stay in L2 across the column loop)
template <int NT>
__device__ void qr_panel_global(const QrItem& it, int j0, PanelShared& sh) {
    const int tid = threadIdx.x;
    const int rows = it.m - j0;
    const int nb = min(QR_NB, it.n - j0);
    double2* base = it.a + j0 + (size_t)it.lda * j0;
    const size_t lda = it.lda;
    qr_panel_columns<NT>(rows, nb, sh, [&](int c, int r) -> double2& { return base[r + lda * c]; });
    __syncthreads();
    // explicit reflectors to v, zeros below the diagonal of R
    for (int e = tid; e < nb * rows; e += NT) {
        const int c = e / rows, r = e % rows;
        const size_t vi = (j0 + r) + (size_t)it.m * (j0 + c);
        double2& ar = base[r + lda * c];
        if (r < c) {
            it.v[vi] = cz();
        } else if (r == c) {
            it.v[vi] = make_double2(1.0, 0.0);
        } else {
            it.v[vi] = ar;
            ar = cz();
        }
    }
    double2* t = it.t + (size_t)(j0 / QR_NB) * QR_NB * QR_NB;
    for (int e = tid; e < QR_NB * QR_NB; e += NT) {
        const int i = e % QR_NB, j = e / QR_NB;
        t[i + QR_NB * j] = (i < nb && j < nb) ? sh.sT[i][j] : cz();
    }
}

// Panel factorisation with the panel in registers, for panels of up to RPT·NT rows:
// thread tid owns rows tid + j·NT (j < RPT) of the (rows × QR_NB) panel, so the sums and
// the trailing update of a column read no memory.  The column index is a template
// parameter (one instantiation per column), so every register index is static.  Same
// arithmetic, in the same order, as qr_panel_columns.
template <int C, int NT, int RPT>
__device__ __forceinline__ void qr_reg_column(double2 (&P)[RPT][QR_NB], int rows, int nb, PanelShared& sh) {
    constexpr int NW = NT / 64;
    const int tid = threadIdx.x;
    double2 sk[QR_NB];
#pragma unroll
    for (int k = 0; k < QR_NB; ++k) sk[k] = cz();
    // (x is 0 outside the rows below the diagonal, and the columns past the panel are 0,
    // so every product is formed without a branch)
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int r = tid + NT * j;
        const double2 x = (r > C && r < rows) ? P[j][C] : cz();
#pragma unroll
        for (int k = 0; k < QR_NB; ++k) {
            if (k < C) sk[k] = cfmac(P[j][k], x, sk[k]);
            else if (k == C) sk[k].x += cabs2(x);
            else sk[k] = cfmac(x, P[j][k], sk[k]);
        }
    }
    if (tid == C) {  // row C (< NT: the thread's first row)
#pragma unroll
        for (int k = 0; k < QR_NB; ++k) sh.sRow[C & 1][k] = P[0][k];
    }
    const double2 tot = block_sum16<NW>(sk, sh.sRed[C & 1]);
    const Reflector h = make_reflector(sh.sRow[C & 1][C], bcast_d(tot.x, C));
    const double2 dl = panel_coef(tid & 15, C, tot, sh.sRow[C & 1][tid & 15], h);
    const double2* dk = panel_dots_publish(dl, sh);
    // v on the owned rows (0 on rows above the diagonal and past the panel)
    double2 vr[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int r = tid + NT * j;
        vr[j] = cz();
        if (r == C) vr[j] = make_double2(1.0, 0.0);
        if (r > C && r < rows) {
            vr[j] = cmul(P[j][C], h.scale);
            P[j][C] = vr[j];
        }
    }
    if (tid == C) P[0][C] = make_double2(h.beta, 0.0);
    if (tid < QR_NB) {
        const int i = tid;
        double2 val = cz();
        if (i < C) {
            double2 acc = cz();
#pragma unroll
            for (int l = 0; l < C; ++l) {
                const double2 t = cfma(sh.sT[i][l], dk[l], acc);
                if (l >= i) acc = t;
            }
            const double2 tc = h.tau;
            val = make_double2(-(tc.x * acc.x - tc.y * acc.y), -(tc.x * acc.y + tc.y * acc.x));
        } else if (i == C) {
            val = h.tau;
        }
        sh.sT[i][C] = val;
    }
    // trailing update a_k -= conj(tau) (v^H a_k) v (v is 0 off the reflector's rows and
    // the dots of the columns past the panel are 0, so no branch is needed)
#pragma unroll
    for (int k = C + 1; k < QR_NB; ++k) {
        const double2 wk = dk[k];
#pragma unroll
        for (int j = 0; j < RPT; ++j) P[j][k] = cfms(wk, vr[j], P[j][k]);
    }
    if constexpr (C + 1 < QR_NB) {
        if (C + 1 < nb) qr_reg_column<C + 1, NT, RPT>(P, rows, nb, sh);
    }
}

// RPT rows per thread: panels of up to RPT·NT rows
template <int NT, int RPT = 2>
__device__ __forceinline__ void qr_panel_regs(const QrItem& it, int j0, PanelShared& sh) {
    const int tid = threadIdx.x;
    const int rows = it.m - j0;
    const int nb = min(QR_NB, it.n - j0);
    const size_t lda = it.lda;
    const double2* a0 = it.a + j0 + lda * j0;
    double2 P[RPT][QR_NB];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int r = tid + NT * j;
#pragma unroll
        for (int k = 0; k < QR_NB; ++k) P[j][k] = ld_or_zero(a0, r + lda * k, r < rows && k < nb);
    }
    qr_reg_column<0, NT, RPT>(P, rows, nb, sh);
    __syncthreads();
    // R (upper part) back to a, the explicit reflectors to v
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        const int r = tid + NT * j;
        if (r >= rows) continue;
#pragma unroll
        for (int k = 0; k < QR_NB; ++k) {
            if (k >= nb) continue;
            const size_t ai = (j0 + r) + lda * (j0 + k);
            const size_t vi = (j0 + r) + (size_t)it.m * (j0 + k);
            if (r <= k) {
                it.a[ai] = P[j][k];
                it.v[vi] = (r == k) ? make_double2(1.0, 0.0) : cz();
            } else {
                it.a[ai] = cz();
                it.v[vi] = P[j][k];
            }
        }
    }
    double2* t = it.t + (size_t)(j0 / QR_NB) * QR_NB * QR_NB;
    for (int e = tid; e < QR_NB * QR_NB; e += NT) {
        const int i = e % QR_NB, j = e / QR_NB;
        t[i + QR_NB * j] = (i < nb && j < nb) ? sh.sT[i][j] : cz();
    }
}

// panels of at most 512 rows (the host sends taller batches to k_qr_panel_tall; the
// global-memory variant is kept out of this kernel and k_qr_step, whose register
// budget it would raise)
__global__ __launch_bounds__(256) void k_qr_panel(const QrItem* __restrict__ items, int j0) {
    __shared__ PanelShared sh;
    const QrItem it = items[blockIdx.x];
    if (j0 >= it.n) return;
    qr_panel_regs<256>(it, j0, sh);
}

// 512 threads: panels of up to 1024 rows in registers, taller ones in global memory
__global__ __launch_bounds__(512) void k_qr_panel_tall(const QrItem* __restrict__ items, int j0) {
    __shared__ PanelShared sh;
    const QrItem it = items[blockIdx.x];
    if (j0 >= it.n) return;
    if (it.m - j0 <= 1024) qr_panel_regs<512>(it, j0, sh);
    else qr_panel_global<512>(it, j0, sh);
}

// Y[j0:, c0 : c0 + tc] <- (I - V T' V^H) Y[j0:, c0 : c0 + tc],  T' = T^H (adjoint) or T,
// V and T of the panel of nb reflectors at j0, on the matrix cores (v_mfma_f64_16x16x4,
// three real products per complex product), operands straight from memory into registers.
// Operand layouts of a 16x16x4 step: lane l gives A[l % 16][l / 16] and
// B[l / 16][l % 16]; result register q of lane l is D[l / 16 + 4 q][l % 16].
//   W = V^H Y: A[i][k] = conj V(r + k, j0 + i), B[k][j] = Y(r + k, c0 + j), four rows a
//     step; wave wv takes steps wv, wv + 4, ..., loads of the next 8 steps in flight
//     while 8 are multiplied.  The four waves' partial sums meet in LDS.
//   W2 = T' W on the vector ALUs (one entry per thread).
//   Y -= V W2 as (V W2)^T = W2^T V^T, so that a lane's results lie along one column
//     of Y: A[j][k] = W2(k, j), B[k][i] = V(r0 + i, j0 + k); one 16-row block per
//     step, wave wv takes blocks wv, wv + 4, ..., two at a time.
template <int NW = 4>  // waves of the workgroup
struct ApplyMfmaShared {
    double2 sPart[NW][QR_NB][QR_TC];
    double2 sW[QR_NB][QR_TC + 1];
    double2 sW2[QR_NB][QR_TC + 1];
    double2 sT[QR_NB][QR_NB + 1];
};
typedef double qr_f64x4 __attribute__((ext_vector_type(4)));
template <int NW = 4>  // waves of the workgroup (4 or 8)
__device__ __forceinline__ void qr_apply_panel_mfma(const QrApplyItem& it, int c0, int tc, int j0, int nb,
                                                    int adjoint, ApplyMfmaShared<NW>& sh) {
    static_assert(QR_TC == QR_NB && QR_NB == 16, "tiles are 16 reflectors × 16 columns");
    constexpr int NT = 64 * NW;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int m = it.m;
    const size_t ldy = it.ldy;
    const double2* t = it.t + (size_t)(j0 / QR_NB) * QR_NB * QR_NB;
    for (int e = tid; e < QR_NB * QR_NB; e += NT) sh.sT[e % QR_NB][e / QR_NB] = gload(t, e);
    // ---- phase 1: W = V^H Y over rows j0 .. m - 1
    const int steps = (m - j0 + 3) / 4;
    const int mine = (steps - wv + NW - 1) / NW;  // this wave's steps: s = wv + NW t, t < mine
    const double2* vcol = it.v + (size_t)m * (j0 + li);
    const double2* ycol = it.y + ldy * (c0 + li);
    const bool vok = li < nb, yok = li < tc;
    constexpr int U = 8;
    double2 a0[U], b0[U], a1[U], b1[U];
    auto fetch = [&](int t0, double2 (&a)[U], double2 (&b)[U]) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int r = j0 + 4 * (wv + NW * (t0 + u)) + lk;
            const bool ok = t0 + u < mine && r < m;
            a[u] = ld_or_zero(vcol, r, ok && vok);
            b[u] = ld_or_zero(ycol, r, ok && yok);
        }
    };
    qr_f64x4 p = {0.0, 0.0, 0.0, 0.0}, q = p, s = p;
    auto mult = [&](const double2 (&a)[U], const double2 (&b)[U]) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            p = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u].x, b[u].x, p, 0, 0, 0);
            q = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u].y, b[u].y, q, 0, 0, 0);
            s = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u].x - a[u].y, b[u].x + b[u].y, s, 0, 0, 0);
        }
    };
    fetch(0, a0, b0);
    for (int t0 = 0; t0 < mine; t0 += 2 * U) {
        fetch(t0 + U, a1, b1);
        mult(a0, b0);
        if (t0 + 2 * U < mine) fetch(t0 + 2 * U, a0, b0);
        mult(a1, b1);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)  // conj(a) b = (P + Q) + i (S - P + Q)
        sh.sPart[wv][lk + 4 * r][li] = make_double2(p[r] + q[r], s[r] - p[r] + q[r]);
    __syncthreads();
    const int wa = (tid / QR_TC) % QR_NB, wc = tid % QR_TC;  // threads 0..255: one entry each
    if (tid < QR_NB * QR_TC) {
        double2 w = sh.sPart[0][wa][wc];
#pragma unroll
        for (int k = 1; k < NW; ++k) w = cadd(w, sh.sPart[k][wa][wc]);
        sh.sW[wa][wc] = w;
    }
    __syncthreads();
    // ---- phase 2: W2 = T' W
    if (tid < QR_NB * QR_TC) {
        double2 w2 = cz();
        for (int b = 0; b < QR_NB; ++b) {
            const double2 tab = adjoint ? cconj(sh.sT[b][wa]) : sh.sT[wa][b];
            w2 = cfma(tab, sh.sW[b][wc], w2);
        }
        sh.sW2[wa][wc] = w2;
    }
    __syncthreads();
    // ---- phase 3: Y -= V W2, 16-row blocks
    double2 wt[4];
#pragma unroll
    for (int k4 = 0; k4 < 4; ++k4) wt[k4] = sh.sW2[4 * k4 + lk][li];
    const int nblk = (m - j0 + 15) / 16;
    for (int blk0 = wv; blk0 < nblk; blk0 += 2 * NW) {
        double2 vv[2][4], yo[2][4];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int blk = blk0 + NW * h;
            const int row = j0 + 16 * blk + li;
            const bool rok = blk < nblk && row < m;
#pragma unroll
            for (int k4 = 0; k4 < 4; ++k4)
                vv[h][k4] = ld_or_zero(it.v, row + (size_t)m * (j0 + 4 * k4 + lk), rok && 4 * k4 + lk < nb);
#pragma unroll
            for (int r = 0; r < 4; ++r)
                yo[h][r] = ld_or_zero(it.y, row + ldy * (c0 + lk + 4 * r), rok && lk + 4 * r < tc);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            qr_f64x4 dp = {0.0, 0.0, 0.0, 0.0}, dq = dp, ds = dp;
#pragma unroll
            for (int k4 = 0; k4 < 4; ++k4) {
                const double2 a = wt[k4], b = vv[h][k4];
                dp = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, dp, 0, 0, 0);
                dq = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.y, dq, 0, 0, 0);
                ds = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x + a.y, b.x + b.y, ds, 0, 0, 0);
            }
            const int blk = blk0 + NW * h;
            const int row = j0 + 16 * blk + li;
            if (blk < nblk && row < m) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {  // a b = (P - Q) + i (S - P - Q)
                    const int c = lk + 4 * r;
                    if (c < tc)
                        gstore(it.y, row + ldy * (c0 + c),
                               csub(yo[h][r], make_double2(dp[r] - dq[r], ds[r] - dp[r] - dq[r])));
                }
            }
        }
    }
}

// one block reflector (panel j0) on one column tile per workgroup
__global__ __launch_bounds__(256) void k_qr_apply(const QrApplyItem* __restrict__ items, const int2* __restrict__ tiles,
                                                  int j0, int adjoint) {
    __shared__ ApplyMfmaShared<4> sh;
    const int2 tt = tiles[blockIdx.x];
    const QrApplyItem it = items[tt.x];
    const int nb = min(QR_NB, it.nrefl - j0);
    if (nb <= 0) return;
    const int c0 = (it.trailing ? j0 + nb : 0) + tt.y * QR_TC;
    const int tc = min(QR_TC, it.ncols - c0);
    if (tc <= 0) return;
    qr_apply_panel_mfma(it, c0, tc, j0, nb, adjoint, sh);
}

// Y <- Q Y with every block reflector of the factorisation, last panel first: the
// column tiles are independent, so one workgroup walks all panels of its tile.
__global__ __launch_bounds__(256) void k_qr_form(const QrApplyItem* __restrict__ items, const int2* __restrict__ tiles) {
    __shared__ ApplyMfmaShared<4> sh;
    const int2 tt = tiles[blockIdx.x];
    const QrApplyItem it = items[tt.x];
    const int c0 = tt.y * QR_TC;
    const int tc = min(QR_TC, it.ncols - c0);
    if (tc <= 0 || it.nrefl <= 0) return;
    for (int j0 = ((it.nrefl - 1) / QR_NB) * QR_NB; j0 >= 0; j0 -= QR_NB) {
        qr_apply_panel_mfma(it, c0, tc, j0, min(QR_NB, it.nrefl - j0), 0, sh);
        __syncthreads();
    }
}

// Look-ahead step of the factorisation: workgroup i < count applies panel j0's block
// reflector to the next panel's columns of item i and then factors that panel; the
// others apply panel j0's block reflector to the trailing tiles beyond it.  The two
// parts touch disjoint columns and disjoint reflector storage, so they run in one
// launch and the panel's latency hides behind the trailing update.
// 512 threads (eight waves, two per SIMD): the panel's column steps run as two waves per
// SIMD with one row per thread, and each trailing tile is applied by eight waves.
constexpr int QR_STEP_NT = 512;
__global__ __launch_bounds__(QR_STEP_NT) __attribute__((amdgpu_waves_per_eu(2))) void k_qr_step(
    const QrItem* __restrict__ items, int count, const QrApplyItem* __restrict__ app, const int2* __restrict__ tiles,
    int j0) {
    __shared__ union {
        PanelShared p;
        ApplyMfmaShared<QR_STEP_NT / 64> a;
    } sh;
    if ((int)blockIdx.x < count) {
        const QrItem it = items[blockIdx.x];
        const int jn = j0 + QR_NB;
        if (jn >= it.n) return;
        const QrApplyItem ai = app[blockIdx.x];  // the same item's reflectors
        qr_apply_panel_mfma<QR_STEP_NT / 64>(ai, jn, min(QR_TC, ai.ncols - jn), j0, QR_NB, 1, sh.a);
        __syncthreads();
        qr_panel_regs<QR_STEP_NT, 1>(it, jn, sh.p);  // at most 512 rows (host routing)
        return;
    }
    const int2 tt = tiles[blockIdx.x - count];
    const QrApplyItem ai = app[tt.x];
    const int nb = min(QR_NB, ai.nrefl - j0);
    if (nb <= 0) return;
    const int c0 = j0 + nb + tt.y * QR_TC;
    const int tc = min(QR_TC, ai.ncols - c0);
    if (tc <= 0) return;
    qr_apply_panel_mfma<QR_STEP_NT / 64>(ai, c0, tc, j0, nb, 1, sh.a);
}

void launch_qr_panel(const QrItem* d_items, int count, int j0, hipStream_t s) {
    if (count <= 0) return;
    k_qr_panel<<<count, 256, 0, s>>>(d_items, j0);
}
void launch_qr_panel_tall(const QrItem* d_items, int count, int j0, hipStream_t s) {
    if (count <= 0) return;
    k_qr_panel_tall<<<count, 512, 0, s>>>(d_items, j0);
}
void launch_qr_apply(const QrApplyItem* d_items, const int2* d_tiles, int ntiles, int j0, int adjoint, hipStream_t s) {
    if (ntiles <= 0) return;
    k_qr_apply<<<ntiles, 256, 0, s>>>(d_items, d_tiles, j0, adjoint);
}
void launch_qr_form(const QrApplyItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s) {
    if (ntiles <= 0) return;
    k_qr_form<<<ntiles, 256, 0, s>>>(d_items, d_tiles);
}
void launch_qr_step(const QrItem* d_items, int count, const QrApplyItem* d_app, const int2* d_tiles, int ntiles,
                    int j0, hipStream_t s) {
    if (count + ntiles <= 0) return;
    k_qr_step<<<count + ntiles, QR_STEP_NT, 0, s>>>(d_items, count, d_app, d_tiles, j0);
}
int qr_tile_cols() { return QR_TC; }

// dst (n × n) = upper triangle of src (ld lds)^H, zeros above the diagonal of dst, for
// a batch (tile map: item, block of 256 entries)
__global__ void k_upper_h_batch(const UpperHItem* __restrict__ items, const int2* __restrict__ tiles) {
    const int2 tt = tiles[blockIdx.x];
    const UpperHItem it = items[tt.x];
    const size_t e = (size_t)tt.y * 256 + threadIdx.x;
    if (e >= (size_t)it.n * it.n) return;
    const int i = (int)(e % it.n), j = (int)(e / it.n);
    it.dst[e] = (i >= j) ? cconj(it.src[j + (size_t)it.lds * i]) : cz();
}
void launch_upper_h_batch(const UpperHItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s) {
    if (ntiles <= 0) return;
    k_upper_h_batch<<<ntiles, 256, 0, s>>>(d_items, d_tiles);
}

// Column order by decreasing norm (the column pivoting of the first pass, taken from
// the initial norms): one 1024-thread workgroup per item, a wave per column for the
// norms, then a rank sort with one thread per column.  A NaN norm sorts last (the
// Jacobi reports the matrix as non-finite).
__global__ __launch_bounds__(1024) void k_pivot_order(const PivotItem* __restrict__ items) {
    __shared__ double sKey[PIVOT_MAXQ];
    const PivotItem it = items[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int j = wv; j < it.q; j += 16) {
        const cplx* col = it.x + (size_t)it.p * j;
        double s = 0.0;
        for (int r = lane; r < it.p; r += 64) s += cabs2(col[r]);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0) sKey[j] = (s == s) ? s : -1.0;
    }
    __syncthreads();
    for (int j = tid; j < it.q; j += 1024) {
        const double k = sKey[j];
        int rank = 0;
        for (int u = 0; u < it.q; ++u) {
            const double ku = sKey[u];
            rank += (ku > k || (ku == k && u < j)) ? 1 : 0;
        }
        it.perm[rank] = j;
        it.perm_inv[j] = rank;
    }
    if (tid == 0) *it.count = it.q;
}
void launch_pivot_order(const PivotItem* d_items, int count, hipStream_t s) {
    if (count > 0) k_pivot_order<<<count, 1024, 0, s>>>(d_items);
}

}  // namespace qh
