// Batched block one-sided (Hestenes) Jacobi SVD in complex FP64, and the
// finalisation kernels: singular values, sort, the `truncate!` rule
// (src/Ansatz/Chain.jl:404-417) and the write-back of the kept factors.
//
// One launch = one round of a round-robin (circle-method) tournament over the
// column blocks of every matrix in the batch.  A workgroup owns one block pair
// (I, J) of one matrix: it forms the Gram matrix G = X_IJ^H X_IJ (BC × BC) from
// HBM, and if any off-diagonal |g_ij| exceeds tol·sqrt(g_ii g_jj) it
// diagonalises G with a cyclic two-sided Jacobi in LDS (FP64) and applies the
// accumulated unitary Q to the columns of X and V:  X_IJ <- X_IJ Q,
// V_IJ <- V_IJ Q.  Convergence is judged on freshly computed inner products, so
// the singular values carry the accuracy of one-sided Jacobi (relative
// accuracy on each σ_j = ||X e_j||).
#include "qh_kernels.h"
#include "qh_complex.h"

typedef double f64x4 __attribute__((ext_vector_type(4)));

namespace qh {

// The tail criterion at the start of an item's own sweep s (after sweep w = s - 1): an
// item of at least JAC_TAIL_NB blocks, past the sorting sweeps, not before
// own[JAC_TAIL_MIN], whose sweep w rotated some but at most tail_switch per mille of its
// block pairs (own[w mod 5]; a sweep without a rotation converges the item in the plain
// loop).
constexpr int JAC_TAIL_NB = 8;
__device__ __forceinline__ bool tail_due(const JacItem& it, int s, int sort_sweeps, int tail_switch) {
    const int w = s - 1;
    if (tail_switch <= 0 || it.nb < JAC_TAIL_NB || w < sort_sweeps || s < it.own[JAC_TAIL_MIN]) return false;
    const long long pairs = (long long)(it.nb - 1) * (it.nb / 2);
    const long long cnt = it.own[w % 5];
    return cnt > 0 && 1000LL * cnt <= tail_switch * pairs;
}
// The round kernels' entry decision, taken by one thread per workgroup: whether the
// workgroup runs (the item is not waiting for the tail passes) and, at the start of the
// item's own sweep `own` (round % R == 0), the tail criterion on its previous sweeps
// (every workgroup of the item reads the same counts and takes the same decision;
// workgroup 0 records it) and the rotation count of the next sweep cleared.
__device__ __forceinline__ int round_gate(const JacItem& it, int round, int own, int R, int pair, int sort_sweeps,
                                          int tail_switch) {
    const int ts = *((volatile int*)it.tail);
    int go = ts != 1;
    if (go && round % R == 0) {
        if (ts == 0 && tail_due(it, own, sort_sweeps, tail_switch)) {
            go = 0;
            if (pair == 0) {
                *it.tail = 1;
                it.own[JAC_TAIL_AT] = own;
            }
        }
        if (go && pair == 0) it.own[(own + 1) % 5] = 0;
    }
    return go;
}
// rr_pair(round + 1, k, nidx) from rr_pair(round, k, nidx)
__device__ __forceinline__ void rr_next(int k, int nidx, int& a, int& b) {
    const int m = nidx - 1;
    if (k == 0) {
        ++a;
    } else {
        a = (a + 1 == m) ? 0 : a + 1;
        b = (b + 1 == m) ? 0 : b + 1;
    }
}


// Rotation of the Hermitian 2×2 problem [[g_aa, g_ab], [conj(g_ab), g_bb]]:
// J = [[c, s e], [-s conj(e), c]] with e = g_ab/|g_ab| and t = s/c the smaller
// root of t^2 + 2 zeta t - 1 = 0, zeta = (g_bb - g_aa)/(2|g_ab|) (Rutishauser's
// formulation); naa / nbb are the new diagonal.  No rotation (s = 0) when the
// pair is already orthogonal, |g_ab|^2 <= tol2·g_aa·g_bb, or a column is below
// the deflation floor.
// With d = g_bb - g_aa and h = sqrt(d^2 + 4|g_ab|^2) (= 2|g_ab| sqrt(1 + zeta^2)):
// |t| = 2|g_ab| / (|d| + h) and c^2 = 1/(1 + t^2) = (|d| + h) / (2h).  The rotation
// is applied through c and se = s e = sign(d) (2c / (|d| + h)) g_ab, and the new
// diagonal through t|g_ab| = sign(d) 2|g_ab|^2 / (|d| + h), so neither |g_ab| nor e
// is formed: two square roots and two divisions, chain sqrt, divide, sqrt.
struct JRot {
    double c = 1.0, sx = 0.0, sy = 0.0, naa = 0.0, nbb = 0.0;  // se = (sx, sy)
};
__device__ __forceinline__ bool rotates(const JRot& j) { return j.sx != 0.0 || j.sy != 0.0; }
__device__ __forceinline__ bool needs_rotation(double gaa, double gbb, double2 gab, double tol2, double gfloor) {
    return gaa > gfloor && gbb > gfloor && cabs2(gab) > tol2 * gaa * gbb;
}
__device__ __forceinline__ JRot jacobi_rot(double gaa, double gbb, double2 gab, double tol2, double gfloor) {
    JRot r;
    r.naa = gaa;
    r.nbb = gbb;
    if (needs_rotation(gaa, gbb, gab, tol2, gfloor)) {
        const double g2 = cabs2(gab);
        const double d = gbb - gaa;
        const double h = sqrt(fma(d, d, 4.0 * g2));
        const double den = fabs(d) + h;
        const double rden = 1.0 / den;
        r.c = sqrt(den / (2.0 * h));
        const double sg = d >= 0.0 ? 2.0 : -2.0;
        const double f = sg * r.c * rden;
        r.sx = f * gab.x;
        r.sy = f * gab.y;
        const double tm = sg * g2 * rden;
        r.naa = gaa - tm;
        r.nbb = gbb + tm;
    }
    return r;
}
// (x_a, x_b) <- (x_a, x_b) J
// (one product and two FMAs per real output)
__device__ __forceinline__ void rot_cols(double2& xa, double2& xb, const JRot& j) {
    const double sx = j.sx, sy = j.sy;  // x_a' = c x_a - conj(se) x_b ; x_b' = se x_a + c x_b
    const double2 na = make_double2(fma(-sy, xb.y, fma(-sx, xb.x, j.c * xa.x)),
                                    fma(sy, xb.x, fma(-sx, xb.y, j.c * xa.y)));
    xb = make_double2(fma(-sy, xa.y, fma(sx, xa.x, j.c * xb.x)), fma(sy, xa.x, fma(sx, xa.y, j.c * xb.y)));
    xa = na;
}
// (y_a; y_b) <- J^H (y_a; y_b)
__device__ __forceinline__ void rot_rows(double2& ya, double2& yb, const JRot& j) {
    const double sx = j.sx, sy = j.sy;  // y_a' = c y_a - se y_b ; y_b' = conj(se) y_a + c y_b
    const double2 na = make_double2(fma(sy, yb.y, fma(-sx, yb.x, j.c * ya.x)),
                                    fma(-sy, yb.x, fma(-sx, yb.y, j.c * ya.y)));
    yb = make_double2(fma(sy, ya.y, fma(sx, ya.x, j.c * yb.x)), fma(-sy, ya.x, fma(sx, ya.y, j.c * yb.y)));
    ya = na;
}

// X_IJ <- X_IJ Q and V_IJ <- V_IJ Q for the 16-row tiles k0 <= k < k1 of X (k < kx =
// ceil(p / 64)) followed by V (the pair's columns in sCol, Q in sQ, both in LDS).
// On the matrix cores (v_mfma_f64_16x16x4), in 16-row tiles: wave mw takes rows
// [64 k + 16 mw, 64 k + 16 mw + 16) of X (k < kx) and then of V, times Q's two
// 16-column tiles.  A complex tile product is three real ones: P1 = Xr Qr,
// P2 = Xi Qi, P3 = (Xr + Xi)(Qr + Qi); re = P1 - P2, im = P3 - P1 - P2.  Operand
// lanes: A[l % 16][l / 16], B[l / 16][l % 16]; result register q of lane l is
// D[l / 16 + 4 q][l % 16].
// A operands straight from HBM: the wave reads and writes only its own rows (16
// consecutive rows of one column per 16 lanes), so the update needs no staging
// and no barrier; Q is read-only in LDS after the eigensolver.  Three tile buffers:
// the loads of the next two tiles are in flight while the matrix cores work on the
// current one.  The loads are unconditional: a row past the end or an
// absent column reads an element in range instead -- absent columns meet zero
// rows of Q (their positions keep e_j), rows past the end are not stored -- so no
// branch holds a load back.
template <int BC>
__device__ __forceinline__ void jac_tiles_update(const JacItem& it, const int* sCol, const double2 (*sQ)[BC + 1],
                                                 int k0, int k1) {
    const int tid = threadIdx.x;
    const int p = it.p;
    const int kx = (p + 63) / 64;
    const int ml = tid & 63, mw = tid >> 6;
    auto tile_load = [&](double2 (&xa)[BC / 4], int k) {
        const bool isv = k >= kx;
        const double2* base = isv ? it.v : it.x;
        const int rows = isv ? it.q : p;
        const int r = min((isv ? k - kx : k) * 64 + 16 * mw + (ml & 15), rows - 1);
#pragma unroll
        for (int st = 0; st < BC / 4; ++st) {
            const int c = sCol[4 * st + (ml >> 4)];
            xa[st] = gload(base, (size_t)r + (size_t)rows * max(c, 0));
        }
    };
    auto tile_update = [&](const double2 (&xa)[BC / 4], int k) {
        f64x4 accP[2], accQ[2], accS[2];
#pragma unroll
        for (int b = 0; b < 2; ++b) accP[b] = accQ[b] = accS[b] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int st = 0; st < BC / 4; ++st) {
            double2 qb[2];
#pragma unroll
            for (int b = 0; b < 2; ++b) qb[b] = sQ[4 * st + (ml >> 4)][16 * b + (ml & 15)];
            const double xs = xa[st].x + xa[st].y;
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const double qs = qb[b].x + qb[b].y;
                accP[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[st].x, qb[b].x, accP[b], 0, 0, 0);
                accQ[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[st].y, qb[b].y, accQ[b], 0, 0, 0);
                accS[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xs, qs, accS[b], 0, 0, 0);
            }
        }
        const bool isv = k >= kx;
        double2* base = isv ? it.v : it.x;
        const int rows = isv ? it.q : p;
        const int r0 = (isv ? k - kx : k) * 64 + 16 * mw + (ml >> 4);
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int gc = sCol[16 * b + (ml & 15)];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int grow = r0 + 4 * q;
                if (gc >= 0 && grow < rows)
                    base[grow + (size_t)rows * gc] =
                        make_double2(accP[b][q] - accQ[b][q], accS[b][q] - accP[b][q] - accQ[b][q]);
            }
        }
    };
    if (k0 >= k1) return;
    double2 xa0[BC / 4], xa1[BC / 4], xa2[BC / 4];
    tile_load(xa0, k0);
    if (k0 + 1 < k1) tile_load(xa1, k0 + 1);
#pragma unroll 1
    for (int k = k0; k < k1; k += 3) {
        if (k + 2 < k1) tile_load(xa2, k + 2);
        tile_update(xa0, k);
        if (k + 1 >= k1) break;
        if (k + 3 < k1) tile_load(xa0, k + 3);
        tile_update(xa1, k + 1);
        if (k + 2 >= k1) break;
        if (k + 4 < k1) tile_load(xa1, k + 4);
        tile_update(xa2, k + 2);
    }
}

// CHECK: the test of every (round, pair) of a sweep in one launch, without rotating (its
// own kernel, so that profiles and the round timings keep the rounds apart)
template <int BC, bool CHECK>
// three waves per SIMD: the MFMA accumulators fit in the VGPRs without spilling
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_jacobi_round(const JacItem* __restrict__ items,
                                                      const int2* __restrict__ tiles, int round_arg,
                                                      const int* __restrict__ round_base, int sort_sweeps,
                                                      int debug, int max_inner, int mode,
                                                      int* __restrict__ flags, const int* __restrict__ flag0,
                                                      int tail_switch, cplx* __restrict__ vq,
                                                      int* __restrict__ vhdr, const int* __restrict__ vmap,
                                                      int nt_pairs, int nv) {
    constexpr int B = BC / 2;
    constexpr int LP = BC + 1;
    constexpr int CH = 32;
    constexpr int S = B / 8;     // Gram sub-tile per thread and block (8 × 8 threads per wave)
    constexpr int NP = BC / 2;   // rotations per inner round
    // the vector-ALU update (BC = 16) stages 128-row × 8-column panels of X through the
    // same LDS as the Gram's 32-row chunks
    constexpr int RB = 128, IC = 8;
    constexpr int SX_ELEMS = (BC * (CH + 1) > IC * (RB + 1)) ? BC * (CH + 1) : IC * (RB + 1);
    __shared__ double2 sG[BC][LP];
    // Q, then the staging area (the MFMA Gram stages its 64-row chunks through both:
    // Q is initialised after the Gram)
    __shared__ double2 sQX[BC * LP + SX_ELEMS];
    double2 (*sQ)[LP] = reinterpret_cast<double2 (*)[LP]>(sQX);
    double2* sXraw = sQX + BC * LP;
    double2 (*sX)[CH + 1] = reinterpret_cast<double2 (*)[CH + 1]>(sXraw);
    __shared__ int sCol[BC];
    __shared__ int sPerm[BC];
    __shared__ int sSlot[BC], sOrd[BC];
    __shared__ double sKey[BC];
    __shared__ int sFlag;
    __shared__ int sGo;

    const int tid = threadIdx.x;
    // the global round of this launch (for JAC_ROUND and JAC_VONLY)
    const int g = round_base ? *round_base + round_arg : round_arg;
    if constexpr (BC == 32 && !CHECK) {
        // V workgroups (with vq: workgroups nt_pairs.. of a round launch, every workgroup
        // of a JAC_VONLY launch): workgroup v applies V_IJ <- V_IJ Q for the Q that pair
        // vmap[nt_pairs + v] stored in the previous round (slot parity (g - 1) & 1) and
        // marks the slot consumed.  The round's own workgroups touch X and the other
        // parity's slots only.
        if (vq != nullptr && (mode == JAC_VONLY || (int)blockIdx.x >= nt_pairs)) {
            const int v = (mode == JAC_VONLY) ? (int)blockIdx.x : (int)blockIdx.x - nt_pairs;
            const int t = vmap[nt_pairs + v];
            const int slot = ((g - 1) & 1) * nv + v;
            if (vhdr[4 * slot] == 0) return;
            const JacItem itv = items[tiles[t].x];
            if (tid < BC) {
                const int c = (tid < B) ? vhdr[4 * slot + 1] * B + tid : vhdr[4 * slot + 2] * B + (tid - B);
                sCol[tid] = (c < itv.q) ? c : -1;
            }
            for (int e = tid; e < BC * BC; e += 256) sQ[e / BC][e % BC] = vq[(size_t)slot * BC * BC + e];
            __syncthreads();  // (every thread has read the flag)
            const int kx = (itv.p + 63) / 64;
            jac_tiles_update<BC>(itv, sCol, sQ, kx, kx + (itv.q + 63) / 64);
            if (tid == 0) vhdr[4 * slot] = 0;
            return;
        }
    }
    const int2 tt = tiles[blockIdx.x];
    const JacItem it = items[tt.x];
    if (*((volatile int*)it.conv)) return;
    const int R = it.nb - 1;
    const double tol = it.tol;
    const double tol2 = tol * tol;
    const double floor_rel2 = it.floor_rel2;
    int round, sort, bi, bj, own = 0;
    if (tid == 0) sGo = 1;
    if (mode == JAC_PAIRS) {
        // an explicit block pair (tile: item, bi | bj << 16), no sorting
        round = -1;
        sort = 0;
        bi = tt.y & 0xffff;
        bj = tt.y >> 16;
    } else if (CHECK) {
        // every (round, pair) of a sweep in one launch (tile: item, round · nb/2 + pair)
        const int np = it.nb / 2;
        round = tt.y / np;
        sort = 0;
        rr_pair(round, tt.y % np, it.nb, bi, bj);
    } else {
        // the round of the tournament: an argument, or the sweep's first round kept in
        // device memory plus the argument (so that a captured sweep replays with
        // advancing rounds); eigenpair sorting during this matrix's own first sweeps
        // (independent of the batch)
        round = g - it.own[JAC_ROUND_OFF];
        own = round / R;
        sort = round < sort_sweeps * R;
        rr_pair(round % R, tt.y, it.nb, bi, bj);
        // (the item's tail state may be set by its workgroup 0 during this launch;
        // counted per own sweep, the decision does not depend on the rest of the batch)
        if (tid == 0) sGo = round_gate(it, round, own, R, tt.y, sort_sweeps, tail_switch);
    }
    if (tid < BC) {
        const int c = (tid < B) ? bi * B + tid : bj * B + (tid - B);
        sCol[tid] = (c < it.q) ? c : -1;
    }
    __syncthreads();
    if (!sGo) return;

    const int p = it.p;
    // columns at or below the rounding floor are frozen (treated as zero)
    const double gfloor = floor_rel2 * (*it.fro2);
    // ---------------- Gram matrix ----------------
    const int lane = tid & 63, wv = tid >> 6;
    const int ti = lane & 7, tj = lane >> 3;
    // G = [[G11, G12], [G12^H, G22]] over the two column blocks: only G11, G12 and
    // G22 are accumulated (3/4 of the products); G21 comes from G12 below
    double2 acc11[S][S], acc12[S][S], acc22[S][S];
#pragma unroll
    for (int a = 0; a < S; ++a)
#pragma unroll
        for (int b = 0; b < S; ++b) acc11[a][b] = acc12[a][b] = acc22[a][b] = cz();

    // chunk loads are software-pipelined: chunk k+1 is fetched into registers
    // while chunk k is reduced from LDS
    constexpr int LPT = BC * CH / 256;  // elements per thread per chunk
    double2 pre[LPT];
    auto fetch = [&](const double2* base, int rows, int row0) {
#pragma unroll
        for (int s2 = 0; s2 < LPT; ++s2) {
            const int e = tid + 256 * s2;
            const int c = e / CH, rr = e % CH;
            const int gc = sCol[c], grow = row0 + rr;
            pre[s2] = ld_or_zero(base, grow + (size_t)rows * gc, gc >= 0 && grow < rows);
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int s2 = 0; s2 < LPT; ++s2) {
            const int e = tid + 256 * s2;
            sX[e / CH][e % CH] = pre[s2];
        }
    };
    if constexpr (BC == 32) {
        // On the matrix cores (v_mfma_f64_16x16x4): a 16-column block is one MFMA tile,
        // and one LDS value per lane gives the operands of X_I^H X_J (A[i][k] = x_i at
        // row k, B[k][j] = x_j at row k).  Three real products per complex block:
        // P1 = ar br, P2 = ai bi, P3 = (ar - ai)(br + bi); Re = P1 + P2, Im = P3 - P1 + P2
        // (the conjugate of A is in the signs).  Rows are staged in 64-row chunks through
        // the storage of Q and the staging area, the next chunk fetched (unconditional
        // loads: a row past the end or an absent column reads an element in range,
        // zeroed when stashed, so no branch holds a load back) while the current one is
        // multiplied.  Wave wv takes rows wv + 4t (t < 16) of each chunk, four per step;
        // result register q of lane l is G[l / 16 + 4 q][l % 16].
        constexpr int CG = 64, LPG = BC * CG / 256;
        static_assert(BC * (CG + 1) <= BC * LP + SX_ELEMS, "a 64-row chunk fits in Q and the staging area");
        double2 (*sXg)[CG + 1] = reinterpret_cast<double2 (*)[CG + 1]>(sQX);
        double2 pg[LPG];
        auto fetch_to = [&](int row0) {
#pragma unroll
            for (int s2 = 0; s2 < LPG; ++s2) {
                const int e = tid + 256 * s2;
                const int c = e / CG, rr = e % CG;
                pg[s2] = gload(it.x, (size_t)min(row0 + rr, p - 1) + (size_t)p * max(sCol[c], 0));
            }
        };
        auto stash_from = [&](int row0) {
#pragma unroll
            for (int s2 = 0; s2 < LPG; ++s2) {
                const int e = tid + 256 * s2;
                const int c = e / CG, rr = e % CG;
                sXg[c][rr] = (sCol[c] >= 0 && row0 + rr < p) ? pg[s2] : cz();
            }
        };
        f64x4 p11 = {0.0, 0.0, 0.0, 0.0}, q11 = p11, s11 = p11, p12 = p11, q12 = p11, s12 = p11;
        f64x4 p22 = p11, q22 = p11, s22 = p11;
        auto reduce_chunk = [&]() {
#pragma unroll
            for (int s4 = 0; s4 < CG / 16; ++s4) {
                const int rr = wv + 4 * (4 * s4 + (lane >> 4));
                const double2 x1 = sXg[lane & 15][rr], x2 = sXg[B + (lane & 15)][rr];
                const double d1 = x1.x - x1.y, e1 = x1.x + x1.y, e2 = x2.x + x2.y, d2 = x2.x - x2.y;
                p11 = __builtin_amdgcn_mfma_f64_16x16x4f64(x1.x, x1.x, p11, 0, 0, 0);
                q11 = __builtin_amdgcn_mfma_f64_16x16x4f64(x1.y, x1.y, q11, 0, 0, 0);
                s11 = __builtin_amdgcn_mfma_f64_16x16x4f64(d1, e1, s11, 0, 0, 0);
                p12 = __builtin_amdgcn_mfma_f64_16x16x4f64(x1.x, x2.x, p12, 0, 0, 0);
                q12 = __builtin_amdgcn_mfma_f64_16x16x4f64(x1.y, x2.y, q12, 0, 0, 0);
                s12 = __builtin_amdgcn_mfma_f64_16x16x4f64(d1, e2, s12, 0, 0, 0);
                p22 = __builtin_amdgcn_mfma_f64_16x16x4f64(x2.x, x2.x, p22, 0, 0, 0);
                q22 = __builtin_amdgcn_mfma_f64_16x16x4f64(x2.y, x2.y, q22, 0, 0, 0);
                s22 = __builtin_amdgcn_mfma_f64_16x16x4f64(d2, e2, s22, 0, 0, 0);
            }
        };
        fetch_to(0);
        for (int row0 = 0; row0 < p; row0 += CG) {
            stash_from(row0);
            __syncthreads();
            if (row0 + CG < p) fetch_to(row0 + CG);
            reduce_chunk();
            __syncthreads();
        }
        for (int w = 0; w < 4; ++w) {
            if (wv == w) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int r = (lane >> 4) + 4 * q, c = lane & 15;
                    const double2 a11 = make_double2(p11[q] + q11[q], s11[q] - p11[q] + q11[q]);
                    const double2 a12 = make_double2(p12[q] + q12[q], s12[q] - p12[q] + q12[q]);
                    const double2 a22 = make_double2(p22[q] + q22[q], s22[q] - p22[q] + q22[q]);
                    sG[r][c] = (w == 0) ? a11 : cadd(sG[r][c], a11);
                    sG[r][B + c] = (w == 0) ? a12 : cadd(sG[r][B + c], a12);
                    sG[B + r][B + c] = (w == 0) ? a22 : cadd(sG[B + r][B + c], a22);
                }
            }
            __syncthreads();
        }
    } else {
        fetch(it.x, p, 0);
        for (int row0 = 0; row0 < p; row0 += CH) {
            stash();
            __syncthreads();
            if (row0 + CH < p) fetch(it.x, p, row0 + CH);
    #pragma unroll 2
            for (int rr = wv; rr < CH; rr += 4) {
                double2 x1i[S], x1j[S], x2i[S], x2j[S];
    #pragma unroll
                for (int a = 0; a < S; ++a) {
                    x1i[a] = sX[ti * S + a][rr];
                    x1j[a] = sX[tj * S + a][rr];
                    x2i[a] = sX[B + ti * S + a][rr];
                    x2j[a] = sX[B + tj * S + a][rr];
                }
    #pragma unroll
                for (int a = 0; a < S; ++a)
    #pragma unroll
                    for (int b = 0; b < S; ++b) {
                        acc11[a][b] = cfmac(x1i[a], x1j[b], acc11[a][b]);
                        acc12[a][b] = cfmac(x1i[a], x2j[b], acc12[a][b]);
                        acc22[a][b] = cfmac(x2i[a], x2j[b], acc22[a][b]);
                    }
            }
            __syncthreads();
        }
        for (int w = 0; w < 4; ++w) {
            if (wv == w) {
    #pragma unroll
                for (int a = 0; a < S; ++a)
    #pragma unroll
                    for (int b = 0; b < S; ++b) {
                        double2& g11 = sG[ti * S + a][tj * S + b];
                        double2& g12 = sG[ti * S + a][B + tj * S + b];
                        double2& g22 = sG[B + ti * S + a][B + tj * S + b];
                        g11 = (w == 0) ? acc11[a][b] : cadd(g11, acc11[a][b]);
                        g12 = (w == 0) ? acc12[a][b] : cadd(g12, acc12[a][b]);
                        g22 = (w == 0) ? acc22[a][b] : cadd(g22, acc22[a][b]);
                    }
            }
            __syncthreads();
        }
    }

    // make G exactly Hermitian so that every test below sees the same value for (i, j)
    for (int e = tid; e < BC * BC; e += 256) {
        const int i = e / BC, j = e % BC;
        if (i > j) sG[i][j] = cconj(sG[j][i]);
        else if (i == j) sG[i][i].y = 0.0;
    }
    __syncthreads();

    // ---------------- convergence test ----------------
    if (tid == 0) sFlag = 0;
    __syncthreads();
    for (int e = tid; e < BC * BC; e += 256) {
        const int i = e / BC, j = e % BC;
        if (i < j) {
            const double gii = sG[i][i].x, gjj = sG[j][j].x;
            const double2 g = sG[i][j];
            if (needs_rotation(gii, gjj, g, tol2, gfloor)) {
                sFlag = 1;
                if (debug)
                    printf("[jr] round %d pair(%d,%d) cols %d %d cos/tol %.6f gii %.17e gjj %.17e\n", round, bi, bj,
                           sCol[i], sCol[j], sqrt(cabs2(g)) / (sqrt(gii) * sqrt(gjj)) / tol, gii, gjj);
            }
        }
    }
    __syncthreads();
    if constexpr (CHECK) {  // the test's outcome only
        if (tid == 0) flags[flag0[tt.x] + tt.y] = sFlag;
        return;
    }
    if (!sFlag) return;

    // the eigensolver and the sort are a latency-bound chain; their waves get issue
    // priority over the co-resident workgroups' streaming Gram and update phases
    __builtin_amdgcn_s_setprio(2);

    // ---------------- EVD of G by cyclic Jacobi ----------------
    // A round applies NP disjoint rotations.  Thread (pb, qb) owns the 2×2 block of G
    // at (row pair pb, column pair qb) and the 2×2 block of Q at rows {2pb, 2pb+1} ×
    // column pair qb.  It derives both rotations from the round's starting G, forms
    // J_pb^H (G_blk J_qb) and Q_blk J_qb in registers and writes them back: G into the
    // other of two buffers (sG and the idle staging area, every entry every round), so
    // one barrier per round separates a round's reads from the next round's writes;
    // Q in place, since rows 2pb, 2pb+1 belong to the NP threads of one wave (whose
    // LDS operations execute in order).
    for (int e = tid; e < BC * BC; e += 256) {
        const int i = e / BC, j = e % BC;
        sQ[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    // one block pair per matrix: solve the block exactly; otherwise an approximate
    // diagonaliser suffices (the outer sweeps re-check every pair on fresh Grams)
    const int inner_cap = (it.nb == 2) ? 12 : max_inner;
    static_assert((NP * NP) % 64 == 0, "block owners must fill whole waves");
    const bool owner = tid < NP * NP;
    const int pb = tid / NP, qb = tid % NP;
    static_assert(SX_ELEMS >= BC * LP, "the staging area holds the second G buffer");
    double2 (*gA)[LP] = sG;                                       // G at the start of the round
    double2 (*gB)[LP] = reinterpret_cast<double2 (*)[LP]>(sXraw);  // G after the round
#pragma unroll 1
    for (int sweep = 0; sweep < inner_cap; ++sweep) {
        if (tid == 0) sFlag = 0;
        __syncthreads();
        // the round's pairs, advanced from round to round without the modulo
        int ap, bp, aq, bq;
        rr_pair(0, pb, BC, ap, bp);
        rr_pair(0, qb, BC, aq, bq);
#pragma unroll 1
        for (int rnd = 0; rnd < BC - 1; ++rnd) {
            if (owner) {
                // each lane solves its column pair; the row pair's rotation comes from
                // the lane of the same wave that owns diagonal block (pb, pb)
                const JRot jq = jacobi_rot(gA[aq][aq].x, gA[bq][bq].x, gA[aq][bq], tol2, gfloor);
                // (whole waves are owners or not, so every lane of the wave shuffles)
                const int src = (pb % (64 / NP)) * NP + pb;
                JRot jp;
                jp.c = __shfl(jq.c, src, 64);
                jp.sx = __shfl(jq.sx, src, 64);
                jp.sy = __shfl(jq.sy, src, 64);
                if (pb == qb) jp = jq;
                double2 g00 = gA[ap][aq], g01 = gA[ap][bq], g10 = gA[bp][aq], g11 = gA[bp][bq];
                if (rotates(jq)) {  // columns:  x_a' = c x_a - s conj(e) x_b ;  x_b' = s e x_a + c x_b
                    double2 q00 = sQ[2 * pb][aq], q01 = sQ[2 * pb][bq];
                    double2 q10 = sQ[2 * pb + 1][aq], q11 = sQ[2 * pb + 1][bq];
                    rot_cols(g00, g01, jq);
                    rot_cols(g10, g11, jq);
                    rot_cols(q00, q01, jq);
                    rot_cols(q10, q11, jq);
                    sQ[2 * pb][aq] = q00; sQ[2 * pb][bq] = q01; sQ[2 * pb + 1][aq] = q10; sQ[2 * pb + 1][bq] = q11;
                }
                if (rotates(jp)) {
                    if (pb == qb) {  // the diagonal block becomes diag(λ_a, λ_b)
                        g00 = make_double2(jp.naa, 0.0);
                        g11 = make_double2(jp.nbb, 0.0);
                        g01 = cz();
                        g10 = cz();
                    } else {  // rows:  y_a' = c y_a - s e y_b ;  y_b' = s conj(e) y_a + c y_b
                        rot_rows(g00, g10, jp);
                        rot_rows(g01, g11, jp);
                    }
                }
                gB[ap][aq] = g00; gB[ap][bq] = g01; gB[bp][aq] = g10; gB[bp][bq] = g11;
                if (pb == qb && rotates(jp)) sFlag = 1;
            }
            rr_next(pb, BC, ap, bp);
            rr_next(qb, BC, aq, bq);
            __syncthreads();
            double2 (*gt)[LP] = gA;
            gA = gB;
            gB = gt;
        }
        const int any = sFlag;
        __syncthreads();
        if (!any) break;
        if (sweep == 0 && tid == 0) {  // counted only if a rotation happened
            atomicAdd(it.rot, 1);
            if (mode == JAC_ROUND) atomicAdd(&it.own[own % 5], 1);
        }
    }

    // ---------------- order eigenpairs: larger eigenvalues to the lower block ----------------
    // (de Rijk-style sorting; valid columns only, invalid positions keep e_j).  Used in the
    // first sweeps only: afterwards the plain cyclic ordering finishes with its usual
    // quadratic convergence (sorting in a parallel ordering can lock into a cycle).
    if (sort) {
    // Rank sort, one thread per column: the valid positions are taken in slot order
    // (the block with the lower index first) and position k of that order receives
    // the column of the k-th largest eigenvalue, ties kept in slot order (a stable
    // descending sort).  Invalid positions keep e_j.
    {
        const int first = (bi < bj) ? 0 : B;
        if (tid < BC) {
            int o = -1;
            double key = 0.0;
            if (sCol[tid] >= 0) {
                const bool in_first = tid >= first && tid < first + B;
                int cnt = 0;
                if (!in_first)
                    for (int u = first; u < first + B; ++u) cnt += (sCol[u] >= 0) ? 1 : 0;
                for (int u = in_first ? first : B - first; u < tid; ++u) cnt += (sCol[u] >= 0) ? 1 : 0;
                o = cnt;
                sSlot[o] = tid;
                const double x = gA[tid][tid].x;  // the eigensolver's final G
                key = (x == x) ? x : -INFINITY;  // a total order even with NaN
            } else {
                sPerm[tid] = tid;
            }
            sOrd[tid] = o;
            sKey[tid] = key;
        }
        __syncthreads();
        if (tid < BC && sOrd[tid] >= 0) {
            const int o = sOrd[tid];
            const double key = sKey[tid];
            int rank = 0;
            for (int u = 0; u < BC; ++u) {
                const int ou = sOrd[u];
                const double ku = sKey[u];
                rank += (ou >= 0 && (ku > key || (ku == key && ou < o))) ? 1 : 0;
            }
            sPerm[sSlot[rank]] = tid;
        }
    }
    __syncthreads();
    for (int e = tid; e < BC * BC; e += 256) {
        const int i = e / BC, j = e % BC;
        sX[i][j] = sQ[i][sPerm[j]];
    }
    __syncthreads();
    for (int e = tid; e < BC * BC; e += 256) {
        const int i = e / BC, j = e % BC;
        sQ[i][j] = sX[i][j];
    }
    __syncthreads();
    }

    __builtin_amdgcn_s_setprio(0);

    // ---------------- X_IJ <- X_IJ Q ,  V_IJ <- V_IJ Q ----------------
    // BC = 32 runs on the matrix cores with its operands loaded directly (below).
    // Otherwise register-blocked on the vector ALUs in 128-row panels: thread (rg, cg)
    // accumulates rows rg + 32a (a < RA) × columns cg·JG + b (b < JG) of a panel, so
    // every LDS value feeds RA or JG complex FMAs; the panel's X is staged IC inner
    // columns at a time (prefetched into registers one stage ahead); the sum runs over
    // i = 0..BC-1 in order.
    constexpr int JG = BC / 8;      // output columns per thread
    constexpr int RA = RB / 32;     // output rows per thread
    constexpr int UPT = IC * RB / 256;
    double2 (*sU)[RB + 1] = reinterpret_cast<double2 (*)[RB + 1]>(sXraw);
    const int rg = tid & 31, cg = tid >> 5;
    double2 upre[UPT];
    auto ufetch = [&](const double2* base, int rows, int row0, int i0) {
#pragma unroll
        for (int s2 = 0; s2 < UPT; ++s2) {
            const int e = tid + 256 * s2;
            const int gc = sCol[i0 + e / RB], grow = row0 + e % RB;
            upre[s2] = ld_or_zero(base, grow + (size_t)rows * gc, gc >= 0 && grow < rows);
        }
    };
    auto ustash = [&]() {
#pragma unroll
        for (int s2 = 0; s2 < UPT; ++s2) {
            const int e = tid + 256 * s2;
            sU[e / RB][e % RB] = upre[s2];
        }
    };
    if constexpr (BC == 32) {
        // X (and V unless V_IJ Q is left to the next launch's V workgroups) on the matrix
        // cores, in 16-row tiles
        const int kx = (p + 63) / 64, kt = kx + (it.q + 63) / 64;
        const int v = (vq != nullptr && mode == JAC_ROUND) ? vmap[blockIdx.x] : -1;
        const bool defer = v >= 0;
        if (defer) {  // Q and its block pair for the V workgroup of the next launch
            const int slot = (g & 1) * nv + v;
            for (int e = tid; e < BC * BC; e += 256) vq[(size_t)slot * BC * BC + e] = sQ[e / BC][e % BC];
            if (tid == 0) {
                vhdr[4 * slot + 1] = bi;
                vhdr[4 * slot + 2] = bj;
                vhdr[4 * slot] = 1;
            }
        }
        jac_tiles_update<BC>(it, sCol, sQ, 0, defer ? kx : kt);
        return;
    }
#pragma unroll 1
    for (int mat = 0; mat < 2; ++mat) {
        double2* base = mat ? it.v : it.x;
        const int rows = mat ? it.q : p;
        ufetch(base, rows, 0, 0);
#pragma unroll 1
        for (int row0 = 0; row0 < rows; row0 += RB) {
            double2 o[RA][JG];
#pragma unroll
            for (int a = 0; a < RA; ++a)
#pragma unroll
                for (int b = 0; b < JG; ++b) o[a][b] = cz();
#pragma unroll 1
            for (int i0 = 0; i0 < BC; i0 += IC) {
                ustash();
                __syncthreads();
                if (i0 + IC < BC) ufetch(base, rows, row0, i0 + IC);
                else if (row0 + RB < rows) ufetch(base, rows, row0 + RB, 0);
#pragma unroll 2
                for (int i = 0; i < IC; ++i) {
                    double2 xv[RA], qv[JG];
#pragma unroll
                    for (int a = 0; a < RA; ++a) xv[a] = sU[i][rg + 32 * a];
#pragma unroll
                    for (int b = 0; b < JG; ++b) qv[b] = sQ[i0 + i][cg * JG + b];
#pragma unroll
                    for (int a = 0; a < RA; ++a)
#pragma unroll
                        for (int b = 0; b < JG; ++b) o[a][b] = cfma(xv[a], qv[b], o[a][b]);
                }
                __syncthreads();
            }
#pragma unroll
            for (int a = 0; a < RA; ++a) {
                const int grow = row0 + rg + 32 * a;
#pragma unroll
                for (int b = 0; b < JG; ++b) {
                    const int gc = sCol[cg * JG + b];
                    if (gc >= 0 && grow < rows) base[grow + (size_t)rows * gc] = o[a][b];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Block width 32: a block pair is 64 columns.  Half as many rounds per sweep as block
// width 16, so each sweep streams X and V through HBM half as often; the price is a
// 64 × 64 eigenproblem per pair (63 inner rounds).  Eight waves; G and Q take 133 KB
// of LDS, so one workgroup per CU.
// ---------------------------------------------------------------------------
// Gram tiles of one wave group (16 × 16 tiles (ti, tj), ti <= tj, of the 4 × 4 grid):
// group 0 takes (0,0) (0,1) (0,2) (0,3) (1,1), group 1 (1,2) (1,3) (2,2) (2,3) (3,3)
template <int GRP>
__device__ __forceinline__ void gram64_step(const double2 (&x)[4], f64x4 (&P)[5], f64x4 (&Q)[5], f64x4 (&S)[5]) {
    constexpr int TI[2][5] = {{0, 0, 0, 0, 1}, {1, 1, 2, 2, 3}};
    constexpr int TJ[2][5] = {{0, 1, 2, 3, 1}, {2, 3, 2, 3, 3}};
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        const double2 a = x[TI[GRP][t]], b = x[TJ[GRP][t]];
        // conj(a)·b by three real products: P = ar br, Q = ai bi, S = (ar - ai)(br + bi)
        P[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, P[t], 0, 0, 0);
        Q[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.y, Q[t], 0, 0, 0);
        S[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x - a.y, b.x + b.y, S[t], 0, 0, 0);
    }
}
template <int GRP>
__device__ __forceinline__ void gram64_store(double2 (*sG)[65], const f64x4 (&P)[5], const f64x4 (&Q)[5],
                                             const f64x4 (&S)[5], bool first, int lane) {
    constexpr int TI[2][5] = {{0, 0, 0, 0, 1}, {1, 1, 2, 2, 3}};
    constexpr int TJ[2][5] = {{0, 1, 2, 3, 1}, {2, 3, 2, 3, 3}};
#pragma unroll
    for (int t = 0; t < 5; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = 16 * TI[GRP][t] + (lane >> 4) + 4 * q, c = 16 * TJ[GRP][t] + (lane & 15);
            const double2 g = make_double2(P[t][q] + Q[t][q], S[t][q] - P[t][q] + Q[t][q]);
            sG[r][c] = first ? g : cadd(sG[r][c], g);
        }
}

__global__ __launch_bounds__(512) void k_jacobi_round64(const JacItem* __restrict__ items,
                                                        const int2* __restrict__ tiles, int round_arg,
                                                        const int* __restrict__ round_base, int sort_sweeps,
                                                        int max_inner, int mode, int tail_switch) {
    constexpr int BC = 64, B = 32, LP = BC + 1, CH = 32, NP = 32, NT = 512;
    __shared__ double2 sG[BC][LP];
    __shared__ double2 sQ[BC][LP];
    __shared__ int sCol[BC];
    __shared__ int sPerm[BC], sSlot[BC], sOrd[BC];
    __shared__ double sKey[BC];
    __shared__ int sFlag;
    __shared__ int sGo;
    // the Gram's 32-row chunks are staged in sQ's storage (Q is initialised afterwards)
    static_assert(BC * (CH + 1) <= BC * LP, "staging fits in sQ");
    double2 (*sX)[CH + 1] = reinterpret_cast<double2 (*)[CH + 1]>(&sQ[0][0]);

    const int tid = threadIdx.x;
    const int2 tt = tiles[blockIdx.x];
    const JacItem it = items[tt.x];
    if (*((volatile int*)it.conv)) return;
    const int R = it.nb - 1;
    const double tol2 = it.tol * it.tol;
    // the modes of k_jacobi_round: a tournament round (with the item's round offset and
    // tail gate), or an explicit block pair of the tail passes (tile: item, bi | bj << 16)
    int round = -1, sort = 0, bi, bj, own = 0;
    if (tid == 0) sGo = 1;
    if (mode == JAC_PAIRS) {
        bi = tt.y & 0xffff;
        bj = tt.y >> 16;
    } else {
        round = (round_base ? *round_base + round_arg : round_arg) - it.own[JAC_ROUND_OFF];
        own = round / R;
        sort = round < sort_sweeps * R;
        rr_pair(round % R, tt.y, it.nb, bi, bj);
        if (tid == 0) sGo = round_gate(it, round, own, R, tt.y, sort_sweeps, tail_switch);
    }
    if (tid < BC) {
        const int c = (tid < B) ? bi * B + tid : bj * B + (tid - B);
        sCol[tid] = (c < it.q) ? c : -1;
    }
    __syncthreads();
    if (!sGo) return;
    const int p = it.p;
    const double gfloor = it.floor_rel2 * (*it.fro2);
    const int lane = tid & 63, wv = tid >> 6;

    // ---------------- Gram matrix (upper 16 × 16 tiles) ----------------
    // Wave group grp = wv / 4 owns five tiles; wave wq = wv % 4 of the group takes rows
    // wq + 4t (t < 8) of each 32-row chunk, four per step; the four partial sums of a
    // tile are added in wave order.
    {
        const int grp = wv >> 2, wq = wv & 3;
        f64x4 P[5], Qa[5], S[5];
#pragma unroll
        for (int t = 0; t < 5; ++t) P[t] = Qa[t] = S[t] = f64x4{0.0, 0.0, 0.0, 0.0};
        constexpr int LPT = BC * CH / NT;
        double2 pre[LPT];
        auto fetch = [&](int row0) {
#pragma unroll
            for (int s2 = 0; s2 < LPT; ++s2) {
                const int e = tid + NT * s2;
                const int c = e / CH, rr = e % CH;
                const int gc = sCol[c], grow = row0 + rr;
                pre[s2] = ld_or_zero(it.x, grow + (size_t)p * gc, gc >= 0 && grow < p);
            }
        };
        fetch(0);
        for (int row0 = 0; row0 < p; row0 += CH) {
#pragma unroll
            for (int s2 = 0; s2 < LPT; ++s2) {
                const int e = tid + NT * s2;
                sX[e / CH][e % CH] = pre[s2];
            }
            __syncthreads();
            if (row0 + CH < p) fetch(row0 + CH);
#pragma unroll
            for (int s4 = 0; s4 < CH / 16; ++s4) {
                const int rr = wq + 4 * (4 * s4 + (lane >> 4));
                double2 x[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) x[t] = sX[16 * t + (lane & 15)][rr];
                if (grp == 0) gram64_step<0>(x, P, Qa, S);
                else gram64_step<1>(x, P, Qa, S);
            }
            __syncthreads();
        }
        for (int w = 0; w < 4; ++w) {
            if (wq == w) {
                if (grp == 0) gram64_store<0>(sG, P, Qa, S, w == 0, lane);
                else gram64_store<1>(sG, P, Qa, S, w == 0, lane);
            }
            __syncthreads();
        }
    }
    // make G exactly Hermitian (the lower tiles were not computed)
    for (int e = tid; e < BC * BC; e += NT) {
        const int i = e / BC, j = e % BC;
        if (i > j) sG[i][j] = cconj(sG[j][i]);
        else if (i == j) sG[i][i].y = 0.0;
    }
    __syncthreads();

    // ---------------- convergence test ----------------
    if (tid == 0) sFlag = 0;
    __syncthreads();
    for (int e = tid; e < BC * BC; e += NT) {
        const int i = e / BC, j = e % BC;
        if (i < j && needs_rotation(sG[i][i].x, sG[j][j].x, sG[i][j], tol2, gfloor)) sFlag = 1;
    }
    __syncthreads();
    if (!sFlag) return;

    __builtin_amdgcn_s_setprio(2);

    // ---------------- EVD of G by cyclic Jacobi ----------------
    // Thread (pb, qb), pb = tid / 32 < 16, owns the 2 × 2 blocks of G at (row pair pb,
    // column pair qb) and (row pair pb + 16, column pair qb), and the matching rows of
    // Q.  Both row rotations come by shuffle from the lanes of the same wave that own
    // the diagonal blocks.  G is updated in place: one barrier after the round's reads,
    // one after its writes; Q's rows belong to one wave each, so Q needs none.
    for (int e = tid; e < BC * BC; e += NT) {
        const int i = e / BC, j = e % BC;
        sQ[i][j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();
    const int inner_cap = (it.nb == 2) ? 12 : max_inner;
    const int pb = tid / NP, qb = tid % NP;
    const int pb2 = pb + 16;
    const int srcA = (pb & 1) * 32 + pb, srcB = srcA + 16;
#pragma unroll 1
    for (int sweep = 0; sweep < inner_cap; ++sweep) {
        if (tid == 0) sFlag = 0;
        __syncthreads();
        int ap, bp, ap2, bp2, aq, bq;
        rr_pair(0, pb, BC, ap, bp);
        rr_pair(0, pb2, BC, ap2, bp2);
        rr_pair(0, qb, BC, aq, bq);
#pragma unroll 1
        for (int rnd = 0; rnd < BC - 1; ++rnd) {
            // the blocks' G and Q entries first, so that their reads overlap the rotation
            double2 a00 = sG[ap][aq], a01 = sG[ap][bq], a10 = sG[bp][aq], a11 = sG[bp][bq];
            double2 b00 = sG[ap2][aq], b01 = sG[ap2][bq], b10 = sG[bp2][aq], b11 = sG[bp2][bq];
            double2 qa[4] = {sQ[2 * pb][aq], sQ[2 * pb][bq], sQ[2 * pb + 1][aq], sQ[2 * pb + 1][bq]};
            double2 qb2[4] = {sQ[2 * pb2][aq], sQ[2 * pb2][bq], sQ[2 * pb2 + 1][aq], sQ[2 * pb2 + 1][bq]};
            const JRot jq = jacobi_rot(sG[aq][aq].x, sG[bq][bq].x, sG[aq][bq], tol2, gfloor);
            JRot jpA, jpB;
            jpA.c = __shfl(jq.c, srcA, 64);
            jpA.sx = __shfl(jq.sx, srcA, 64);
            jpA.sy = __shfl(jq.sy, srcA, 64);
            jpB.c = __shfl(jq.c, srcB, 64);
            jpB.sx = __shfl(jq.sx, srcB, 64);
            jpB.sy = __shfl(jq.sy, srcB, 64);
            if (pb == qb) jpA = jq;
            if (pb2 == qb) jpB = jq;
            __syncthreads();
            if (rotates(jq)) {
                rot_cols(a00, a01, jq);
                rot_cols(a10, a11, jq);
                rot_cols(b00, b01, jq);
                rot_cols(b10, b11, jq);
                rot_cols(qa[0], qa[1], jq);
                rot_cols(qa[2], qa[3], jq);
                rot_cols(qb2[0], qb2[1], jq);
                rot_cols(qb2[2], qb2[3], jq);
                sQ[2 * pb][aq] = qa[0]; sQ[2 * pb][bq] = qa[1]; sQ[2 * pb + 1][aq] = qa[2]; sQ[2 * pb + 1][bq] = qa[3];
                sQ[2 * pb2][aq] = qb2[0]; sQ[2 * pb2][bq] = qb2[1];
                sQ[2 * pb2 + 1][aq] = qb2[2]; sQ[2 * pb2 + 1][bq] = qb2[3];
            }
            if (rotates(jpA)) {
                if (pb == qb) {
                    a00 = make_double2(jpA.naa, 0.0);
                    a11 = make_double2(jpA.nbb, 0.0);
                    a01 = cz();
                    a10 = cz();
                } else {
                    rot_rows(a00, a10, jpA);
                    rot_rows(a01, a11, jpA);
                }
            }
            if (rotates(jpB)) {
                if (pb2 == qb) {
                    b00 = make_double2(jpB.naa, 0.0);
                    b11 = make_double2(jpB.nbb, 0.0);
                    b01 = cz();
                    b10 = cz();
                } else {
                    rot_rows(b00, b10, jpB);
                    rot_rows(b01, b11, jpB);
                }
            }
            sG[ap][aq] = a00; sG[ap][bq] = a01; sG[bp][aq] = a10; sG[bp][bq] = a11;
            sG[ap2][aq] = b00; sG[ap2][bq] = b01; sG[bp2][aq] = b10; sG[bp2][bq] = b11;
            if ((pb == qb && rotates(jpA)) || (pb2 == qb && rotates(jpB))) sFlag = 1;
            rr_next(pb, BC, ap, bp);
            rr_next(pb2, BC, ap2, bp2);
            rr_next(qb, BC, aq, bq);
            __syncthreads();
        }
        const int any = sFlag;
        __syncthreads();
        if (!any) break;
        if (sweep == 0 && tid == 0) {  // counted only if a rotation happened
            atomicAdd(it.rot, 1);
            if (mode == JAC_ROUND) atomicAdd(&it.own[own % 5], 1);
        }
    }

    // ---------------- order eigenpairs (first sweeps only; as for block width 16) ----------------
    if (sort) {
        const int first = (bi < bj) ? 0 : B;
        if (tid < BC) {
            int o = -1;
            double key = 0.0;
            if (sCol[tid] >= 0) {
                const bool in_first = tid >= first && tid < first + B;
                int cnt = 0;
                if (!in_first)
                    for (int u = first; u < first + B; ++u) cnt += (sCol[u] >= 0) ? 1 : 0;
                for (int u = in_first ? first : B - first; u < tid; ++u) cnt += (sCol[u] >= 0) ? 1 : 0;
                o = cnt;
                sSlot[o] = tid;
                const double x = sG[tid][tid].x;
                key = (x == x) ? x : -INFINITY;
            } else {
                sPerm[tid] = tid;
            }
            sOrd[tid] = o;
            sKey[tid] = key;
        }
        __syncthreads();
        if (tid < BC && sOrd[tid] >= 0) {
            const int o = sOrd[tid];
            const double key = sKey[tid];
            int rank = 0;
            for (int u = 0; u < BC; ++u) {
                const int ou = sOrd[u];
                const double ku = sKey[u];
                rank += (ou >= 0 && (ku > key || (ku == key && ou < o))) ? 1 : 0;
            }
            sPerm[sSlot[rank]] = tid;
        }
        __syncthreads();
        for (int e = tid; e < BC * BC; e += NT) {
            const int i = e / BC, j = e % BC;
            sG[i][j] = sQ[i][sPerm[j]];
        }
        __syncthreads();
        for (int e = tid; e < BC * BC; e += NT) {
            const int i = e / BC, j = e % BC;
            sQ[i][j] = sG[i][j];
        }
        __syncthreads();
    }

    __builtin_amdgcn_s_setprio(0);

    // ---------------- X_IJ <- X_IJ Q ,  V_IJ <- V_IJ Q  (matrix cores) ----------------
    // 128-row panels of X and then of V: wave wv owns panel rows [16 wv, 16 wv + 16)
    // times Q's four 16-column tiles; operands straight from HBM (each wave reads and
    // writes only its own rows), the next panel's sixteen loads issued before the
    // current panel's products (unconditional, in range: see the width-16 update);
    // three real products per complex tile: re = P1 - P2, im = P3 - P1 - P2.
    const int kx = (p + 127) / 128, kt = kx + (it.q + 127) / 128;
    auto panel_load = [&](double2 (&xall)[BC / 4], int k) {
        const bool isv = k >= kx;
        const double2* base = isv ? it.v : it.x;
        const int rows = isv ? it.q : p;
        const int rw = min((isv ? k - kx : k) * 128 + 16 * wv + (lane & 15), rows - 1);
#pragma unroll
        for (int st = 0; st < BC / 4; ++st) {
            const int c = sCol[4 * st + (lane >> 4)];
            xall[st] = gload(base, (size_t)rw + (size_t)rows * max(c, 0));
        }
    };
    auto panel_update = [&](const double2 (&xall)[BC / 4], int k) {
        f64x4 accP[4], accQ[4], accS[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) accP[b] = accQ[b] = accS[b] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int st = 0; st < BC / 4; ++st) {
            // keep each step's Q reads in their step (the loads stay in flight)
            __builtin_amdgcn_sched_barrier(0);
            const double2 xa = xall[st];
            const double xs = xa.x + xa.y;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const double2 qv = sQ[4 * st + (lane >> 4)][16 * b + (lane & 15)];
                accP[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa.x, qv.x, accP[b], 0, 0, 0);
                accQ[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa.y, qv.y, accQ[b], 0, 0, 0);
                accS[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xs, qv.x + qv.y, accS[b], 0, 0, 0);
            }
        }
        const bool isv = k >= kx;
        double2* base = isv ? it.v : it.x;
        const int rows = isv ? it.q : p;
        const int r0 = (isv ? k - kx : k) * 128 + 16 * wv + (lane >> 4);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int gc = sCol[16 * b + (lane & 15)];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int grow = r0 + 4 * q;
                if (gc >= 0 && grow < rows)
                    base[grow + (size_t)rows * gc] =
                        make_double2(accP[b][q] - accQ[b][q], accS[b][q] - accP[b][q] - accQ[b][q]);
            }
        }
    };
    double2 xa0[BC / 4], xa1[BC / 4];
    panel_load(xa0, 0);
#pragma unroll 1
    for (int k = 0; k < kt; k += 2) {
        if (k + 1 < kt) panel_load(xa1, k + 1);
        panel_update(xa0, k);
        if (k + 1 >= kt) break;
        if (k + 2 < kt) panel_load(xa0, k + 2);
        panel_update(xa1, k + 1);
    }
}

void launch_jacobi_round(const JacItem* d_items, const int2* d_tiles, int ntiles, int round,
                         const int* round_base, int block, int sort_sweeps, int debug, int max_inner, int tail_switch,
                         cplx* vq, int* vhdr, const int* vmap, int nv, hipStream_t s) {
    if (ntiles <= 0) return;
    if (block == 8)
        k_jacobi_round<16, false><<<ntiles, 256, 0, s>>>(d_items, d_tiles, round, round_base, sort_sweeps, debug,
                                                         max_inner, JAC_ROUND, nullptr, nullptr,
                                                         tail_switch, nullptr, nullptr, nullptr, ntiles, 0);
    else if (block == 32)
        k_jacobi_round64<<<ntiles, 512, 0, s>>>(d_items, d_tiles, round, round_base, sort_sweeps, max_inner,
                                                JAC_ROUND, tail_switch);
    else  // with vq: the pairs' workgroups, then the V workgroups
        k_jacobi_round<32, false><<<vq ? ntiles + nv : ntiles, 256, 0, s>>>(
            d_items, d_tiles, round, round_base, sort_sweeps, debug, max_inner, JAC_ROUND, nullptr, nullptr,
            tail_switch, vq, vhdr, vmap, ntiles, nv);
}
void launch_jacobi_vonly(const JacItem* d_items, const int2* d_tiles, int ntiles, int round, const int* round_base,
                         cplx* vq, int* vhdr, const int* vmap, int nv, hipStream_t s) {
    if (nv <= 0) return;
    k_jacobi_round<32, false><<<nv, 256, 0, s>>>(d_items, d_tiles, round, round_base, 0, 0, 1, JAC_VONLY, nullptr,
                                                 nullptr, 0, vq, vhdr, vmap, ntiles, nv);
}
void launch_jacobi_check(const JacItem* d_items, const int2* d_tiles, int ntiles, int* d_flags, const int* d_flag0,
                         int block, hipStream_t s) {
    if (ntiles <= 0) return;
    if (block == 8)
        k_jacobi_round<16, true><<<ntiles, 256, 0, s>>>(d_items, d_tiles, 0, nullptr, 0, 0, 1, JAC_CHECK, d_flags, d_flag0,
                                                        0, nullptr, nullptr, nullptr, ntiles, 0);
    else
        k_jacobi_round<32, true><<<ntiles, 256, 0, s>>>(d_items, d_tiles, 0, nullptr, 0, 0, 1, JAC_CHECK, d_flags, d_flag0,
                                                        0, nullptr, nullptr, nullptr, ntiles, 0);
}
// Convergence test of whole items without rotating (block width 16 or 32): G = X^H X in
// 64 × 64 tiles of the upper triangle of 16-column blocks, one workgroup per tile, on the
// matrix cores (v_mfma_f64_16x16x4, three real products per complex block, as in the
// rounds' Gram); wave w holds the four 16 × 16 sub-tiles of row w of the tile, i.e. the
// 16-column block pairs (4·ti + w, 4·tj + b).  The column norms come from the same LDS
// chunks.  flags[first[item] + I·nb + J] is set to 1 where a column pair of the item's
// blocks I <= J (bsub = block / 16 sixteen-column blocks each) fails the rotation test
// (for I = J the pairs inside the block); the table starts at zero.  Rows are streamed in
// chunks of 16, the next chunk fetched while the current one is multiplied.
constexpr int JCT = 64, JCR = 16;
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_jacobi_check_tiles(const JacItem* __restrict__ items,
                                                            const int4* __restrict__ tiles, int* __restrict__ flags,
                                                            const int* __restrict__ first, int bsub) {
    __shared__ double2 sX[2 * JCT][JCR + 1];
    __shared__ double sN[2 * JCT];
    const int4 tt = tiles[blockIdx.x];
    const JacItem it = items[tt.x];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int p = it.p, q = it.q;
    const int c0 = tt.y * JCT, c1 = tt.z * JCT;
    // thread tid stages rows e % JCR of column e / JCR, e = tid + 256 s (side I for the
    // columns below JCT, side J above)
    constexpr int LPT = 2 * JCT * JCR / 256;
    double2 pre[LPT];
    auto fetch = [&](int row0) {
#pragma unroll
        for (int s2 = 0; s2 < LPT; ++s2) {
            const int e = tid + 256 * s2;
            const int cl = e / JCR, r = row0 + e % JCR;
            const int col = (cl < JCT) ? c0 + cl : c1 + cl - JCT;
            pre[s2] = ld_or_zero(it.x, (size_t)min(r, p - 1) + (size_t)p * min(col, q - 1), r < p && col < q);
        }
    };
    f64x4 P[4], Q[4], S[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) P[b] = Q[b] = S[b] = f64x4{0.0, 0.0, 0.0, 0.0};
    double nrm = 0.0;  // thread tid < 2·JCT: |x|^2 of column tid of the tile pair
    fetch(0);
    for (int row0 = 0; row0 < p; row0 += JCR) {
#pragma unroll
        for (int s2 = 0; s2 < LPT; ++s2) {
            const int e = tid + 256 * s2;
            sX[e / JCR][e % JCR] = pre[s2];
        }
        __syncthreads();
        if (row0 + JCR < p) fetch(row0 + JCR);
        if (tid < 2 * JCT) {
#pragma unroll
            for (int r = 0; r < JCR; ++r) nrm += cabs2(sX[tid][r]);
        }
#pragma unroll
        for (int s4 = 0; s4 < JCR / 4; ++s4) {
            const int r = 4 * s4 + (lane >> 4);
            const double2 xa = sX[16 * wv + (lane & 15)][r];
            const double da = xa.x - xa.y;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const double2 xb = sX[JCT + 16 * b + (lane & 15)][r];
                P[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa.x, xb.x, P[b], 0, 0, 0);
                Q[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa.y, xb.y, Q[b], 0, 0, 0);
                S[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(da, xb.x + xb.y, S[b], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    if (tid < 2 * JCT) sN[tid] = nrm;
    __syncthreads();
    const double tol2 = it.tol * it.tol;
    const double gfloor = it.floor_rel2 * (*it.fro2);
    const bool diag = tt.y == tt.z;
    const int n16 = it.nb * bsub;  // 16-column blocks of the item
    const int I = 4 * tt.y + wv;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int J = 4 * tt.z + b;
        if (I >= n16 || J >= n16 || (diag && b < wv)) continue;  // wave-uniform
        bool bad = false;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = (lane >> 4) + 4 * r, j = lane & 15;  // result register r of the lane
            if (diag && b == wv && i >= j) continue;
            const double2 g = make_double2(P[b][r] + Q[b][r], S[b][r] - P[b][r] + Q[b][r]);
            bad = bad || needs_rotation(sN[16 * wv + i], sN[JCT + 16 * b + j], g, tol2, gfloor);
        }
        const bool any = __ballot(bad) != 0;
        if (lane == 0 && any) flags[first[tt.x] + (I / bsub) * it.nb + J / bsub] = 1;
    }
}
void launch_jacobi_check_tiles(const JacItem* d_items, const int4* d_tiles, int ntiles, int* d_flags,
                               const int* d_first, int block, hipStream_t s) {
    if (ntiles <= 0) return;
    k_jacobi_check_tiles<<<ntiles, 256, 0, s>>>(d_items, d_tiles, d_flags, d_first, block / 16);
}
void launch_jacobi_pairs(const JacItem* d_items, const int2* d_tiles, int ntiles, int block, int max_inner,
                         hipStream_t s) {
    if (ntiles <= 0) return;
    if (block == 32)
        k_jacobi_round64<<<ntiles, 512, 0, s>>>(d_items, d_tiles, 0, nullptr, 0, max_inner, JAC_PAIRS, 0);
    else if (block == 8)
        k_jacobi_round<16, false><<<ntiles, 256, 0, s>>>(d_items, d_tiles, 0, nullptr, 0, 0, max_inner, JAC_PAIRS,
                                                         nullptr, nullptr, 0, nullptr, nullptr, nullptr, ntiles, 0);
    else
        k_jacobi_round<32, false><<<ntiles, 256, 0, s>>>(d_items, d_tiles, 0, nullptr, 0, 0, max_inner, JAC_PAIRS,
                                                         nullptr, nullptr, 0, nullptr, nullptr, nullptr, ntiles, 0);
}

// ||X||_F^2 of every item, deterministic: each workgroup reduces one chunk into
// partial[tile]; a second pass adds an item's chunks in order.  The sum, and so the
// deflation floor, does not depend on scheduling.
__global__ __launch_bounds__(256) void k_frob2(const JacItem* __restrict__ items, const int2* __restrict__ tiles,
                                               double* __restrict__ partial) {
    __shared__ double sRed[4];
    const int2 tt = tiles[blockIdx.x];
    const JacItem it = items[tt.x];
    const size_t total = (size_t)it.p * it.q;
    const size_t beg = (size_t)tt.y * FROB_CHUNK;
    const size_t end = beg + FROB_CHUNK < total ? beg + FROB_CHUNK : total;
    double s = 0.0;
    for (size_t e = beg + threadIdx.x; e < end; e += 256) s += cabs2(it.x[e]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sRed[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (sRed[0] + sRed[1]) + (sRed[2] + sRed[3]);
}
// item i owns partial[first[i] .. first[i+1])
__global__ void k_frob2_sum(const JacItem* __restrict__ items, int count, const double* __restrict__ partial,
                            const int* __restrict__ first) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    double s = 0.0;
    for (int k = first[i]; k < first[i + 1]; ++k) s += partial[k];
    *items[i].fro2 = s;
    if (!isfinite(s)) *items[i].conv = -1;  // NaN / Inf input: never rotated, reported by the host
}
void launch_frob2(const JacItem* d_items, int count, const int2* d_tiles, int ntiles, double* partial,
                  const int* first, hipStream_t s) {
    if (count <= 0) return;
    if (ntiles > 0) k_frob2<<<ntiles, 256, 0, s>>>(d_items, d_tiles, partial);
    k_frob2_sum<<<(count + 255) / 256, 256, 0, s>>>(d_items, count, partial, first);
}

// V <- I for every item (grid: items × column tiles)
__global__ void k_set_identity(const JacItem* __restrict__ items, const int2* __restrict__ tiles) {
    const int2 tt = tiles[blockIdx.x];
    const JacItem it = items[tt.x];
    const int col = tt.y;
    if (col >= it.q) return;
    for (int r = threadIdx.x; r < it.q; r += blockDim.x)
        it.v[r + (size_t)it.q * col] = make_double2(r == col ? 1.0 : 0.0, 0.0);
    if (threadIdx.x == 0) {
        *it.rot = 0;
        *it.conv = 0;
        *it.tail = 0;
        for (int k = 0; k < JAC_OWN_INTS; ++k) it.own[k] = 0;
    }
}
void launch_set_identity_items(const JacItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s) {
    if (ntiles <= 0) return;
    k_set_identity<<<ntiles, 256, 0, s>>>(d_items, d_tiles);
}

// After a window of rounds: an item with no rotation is converged.  remaining[0]
// receives the number of items still rotating, remaining[1] the number with
// non-finite entries.
__global__ void k_conv_update(const JacItem* __restrict__ items, int count, int* remaining, int* round_base,
                              int rounds, int first_round, int sort_sweeps, int tail_switch) {
    __shared__ int sRem, sBad, sTail;
    if (threadIdx.x == 0) sRem = sBad = sTail = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
        const JacItem it = items[i];
        const int cv = *it.conv;
        // the window's rotation count and the tail state it ran under, for the host's
        // accounting (read back with `remaining`)
        remaining[4 + 2 * i] = *it.rot;
        remaining[4 + 2 * i + 1] = *it.tail;
        if (cv == 0) {
            // an item whose own sweep ends with the window takes the round kernel's tail
            // decision now, so that the next window does not start with it
            const int R = it.nb - 1, end = first_round + rounds - it.own[JAC_ROUND_OFF];
            if (*it.tail == 0 && *it.rot > 0 && end % R == 0 && tail_due(it, end / R, sort_sweeps, tail_switch)) {
                *it.tail = 1;
                it.own[JAC_TAIL_AT] = end / R;
            }
            if (*it.tail == 1) atomicAdd(&sTail, 1);  // waits for the tail passes
            else if (*it.rot == 0) *it.conv = 1;
            else atomicAdd(&sRem, 1);
        } else if (cv < 0) {
            atomicAdd(&sBad, 1);
        }
        *it.rot = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        remaining[0] = sRem;
        remaining[1] = sBad;
        remaining[3] = sTail;
        if (round_base) *round_base += rounds;  // the next sweep's first round
    }
}
void launch_conv_update(const JacItem* d_items, int count, int* remaining, int* round_base, int rounds,
                        int first_round, int sort_sweeps, int tail_switch, hipStream_t s) {
    k_conv_update<<<1, 256, 0, s>>>(d_items, count, remaining, round_base, rounds, first_round, sort_sweeps,
                                    tail_switch);
}

// ---------------------------------------------------------------------------
// singular values: σ_j = ||X[:, j]||  (one wave per column)
// ---------------------------------------------------------------------------
__global__ void k_colnorms(const SelectItem* __restrict__ items, const int2* __restrict__ tiles) {
    const int2 tt = tiles[blockIdx.x];
    const SelectItem it = items[tt.x];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int col = tt.y * 4 + wv;
    if (col >= it.q) return;
    const double2* x = it.x + (size_t)it.p * col;
    double s = 0.0;
    for (int r = lane; r < it.p; r += 64) s += cabs2(x[r]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) it.sigma_sorted[col] = sqrt(s);
}

// sort descending, `truncate!` rule, renormalisation, discarded weight
constexpr int SELECT_MAXQ = 4096;
__global__ __launch_bounds__(1024) void k_select(const SelectItem* __restrict__ items) {
    __shared__ double sS[SELECT_MAXQ];
    __shared__ int sI[SELECT_MAXQ];
    __shared__ double sRed[32];
    __shared__ int sKept;
    const SelectItem it = items[blockIdx.x];
    const int tid = threadIdx.x, q = it.q;
    int n2 = 1;
    while (n2 < q) n2 <<= 1;
    for (int i = tid; i < n2; i += blockDim.x) {
        sS[i] = (i < q) ? it.sigma_sorted[i] : -1.0;
        sI[i] = (i < q) ? i : 0x7fffffff;
    }
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = tid; i < n2; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const double si = sS[i], sj = sS[ixj];
                    const int ii = sI[i], ij = sI[ixj];
                    // "i before ixj" in descending order
                    const bool i_first = (si > sj) || (si == sj && ii < ij);
                    const bool desc = ((i & k) == 0);
                    if (desc != i_first) {
                        sS[i] = sj; sS[ixj] = si;
                        sI[i] = ij; sI[ixj] = ii;
                    }
                }
            }
            __syncthreads();
        }
    }
    if (tid == 0) {
        int lim = q;
        int kept = q;
        if (it.truncate) {
            if (it.maxdim >= 0 && it.maxdim < lim) lim = it.maxdim;
            kept = 0;
            while (kept < lim && fabs(sS[kept]) > it.threshold) ++kept;
        }
        sKept = kept;
    }
    __syncthreads();
    const int kept = sKept;
    // discarded weight and kept norm
    double dsc = 0.0, kn = 0.0;
    for (int i = tid; i < q; i += blockDim.x) {
        const double v = sS[i] * sS[i];
        if (i < kept) kn += v; else dsc += v;
    }
    for (int off = 32; off > 0; off >>= 1) {
        dsc += __shfl_xor(dsc, off, 64);
        kn += __shfl_xor(kn, off, 64);
    }
    const int lane = tid & 63, wv = tid >> 6;
    if (lane == 0) { sRed[wv] = dsc; sRed[16 + wv] = kn; }
    __syncthreads();
    if (tid == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { a += sRed[w]; b += sRed[16 + w]; }
        sRed[0] = a; sRed[16] = b;
        *it.kept = kept;
        *it.discarded = sqrt(a);
    }
    __syncthreads();
    const double scale = (it.renormalize && sRed[16] > 0.0) ? 1.0 / sqrt(sRed[16]) : 1.0;
    for (int i = tid; i < q; i += blockDim.x) {
        it.sigma_sorted[i] = sS[i];
        it.perm[i] = sI[i];
        if (i < kept) it.lam_out[i] = sS[i] * scale;
    }
}

void launch_select_items(const SelectItem* d_items, const int2* d_norm_tiles, int n_norm_tiles, int count,
                         hipStream_t s) {
    if (n_norm_tiles > 0) k_colnorms<<<n_norm_tiles, 256, 0, s>>>(d_items, d_norm_tiles);
    if (count > 0) k_select<<<count, 1024, 0, s>>>(d_items);
}

// ---------------------------------------------------------------------------
// orthonormal completion of kept columns below the deflation floor
// ---------------------------------------------------------------------------
__device__ __forceinline__ double splitmix_unit(unsigned long long k) {
    k += 0x9e3779b97f4a7c15ull;
    k = (k ^ (k >> 30)) * 0xbf58476d1ce4e5b9ull;
    k = (k ^ (k >> 27)) * 0x94d049bb133111ebull;
    k ^= k >> 31;
    return (double)(k >> 11) * 0x1.0p-52 - 1.0;  // uniform in [-1, 1)
}

constexpr int COMPLETE_MAXQ = SELECT_MAXQ;
// The kept columns at or below the deflation floor (sorted positions [first, kept);
// σ sorted decreasing, so first = the number of kept σ² above the floor) are replaced
// by an orthonormal completion: deterministic pseudo-random starts, two passes of
// classical Gram-Schmidt against the columns above the floor (one workgroup per
// frozen column, all in parallel); the host then orthonormalises them among
// themselves by a Householder QR.  Sorted column i is x[:, perm[i]] / s_i with
// s_i = σ_i, or 1 where σ_i = 0.
__device__ __forceinline__ int complete_first(const CompleteItem& it, int kept, double floor2) {
    int lo = 0, hi = kept;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const double s = it.sigma_sorted[mid];
        if (s * s > floor2) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ const double2* complete_col(const CompleteItem& it, int i) {
    return it.x + (size_t)it.p * it.perm[i];
}
__device__ __forceinline__ double complete_inv(const CompleteItem& it, int i) {
    const double s = it.sigma_sorted[i];
    return s > 0.0 ? 1.0 / s : 1.0;
}
// sDot[i - i0] = <x_i, z> / s_i² for i in [i0, i1): a wave per column, four columns
// per wave in flight, lanes striding the rows
__device__ __forceinline__ void complete_dots(const CompleteItem& it, const double2* z, int i0, int i1,
                                              double2* sDot) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = i0 + wv; i < i1; i += 16) {
        double2 acc[4] = {cz(), cz(), cz(), cz()};
        const double2* cc[4];
        bool ok[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            ok[u] = i + 4 * u < i1;
            cc[u] = complete_col(it, ok[u] ? i + 4 * u : i);
        }
        for (int r = lane; r < it.p; r += 64) {
            const double2 zr = z[r];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (ok[u]) acc[u] = cfmac(cc[u][r], zr, acc[u]);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                acc[u].x += __shfl_xor(acc[u].x, off, 64);
                acc[u].y += __shfl_xor(acc[u].y, off, 64);
            }
        if (lane == 0)
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (ok[u]) {
                    const double w = complete_inv(it, i + 4 * u);
                    sDot[i + 4 * u - i0] = cscale(acc[u], w * w);
                }
    }
}
// z <- z - Σ_{i in [i0, i1)} sDot[i - i0] x_i, a thread per row
__device__ __forceinline__ void complete_update(const CompleteItem& it, double2* z, int i0, int i1,
                                                const double2* sDot) {
    for (int r = threadIdx.x; r < it.p; r += blockDim.x) {
        double2 v = z[r];
        int i = i0;
        for (; i + 4 <= i1; i += 4) {
            const double2 x0 = complete_col(it, i)[r], x1 = complete_col(it, i + 1)[r];
            const double2 x2 = complete_col(it, i + 2)[r], x3 = complete_col(it, i + 3)[r];
            v = csub(v, cmul(sDot[i - i0], x0));
            v = csub(v, cmul(sDot[i + 1 - i0], x1));
            v = csub(v, cmul(sDot[i + 2 - i0], x2));
            v = csub(v, cmul(sDot[i + 3 - i0], x3));
        }
        for (; i < i1; ++i) v = csub(v, cmul(sDot[i - i0], complete_col(it, i)[r]));
        z[r] = v;
    }
}

// phase 1, workgroup (item, j): a random start for frozen column j, orthogonalised
// twice against the columns above the floor
__global__ __launch_bounds__(256) void k_complete_project(const CompleteItem* __restrict__ items,
                                                          const int2* __restrict__ tiles) {
    __shared__ double2 sDot[COMPLETE_MAXQ];
    const int2 tt = tiles[blockIdx.x];
    const CompleteItem it = items[tt.x];
    const int j = tt.y;
    const int kept = *it.kept;
    if (j >= kept) return;
    const double floor2 = it.floor_rel2 * (*it.fro2);
    const double sj = it.sigma_sorted[j];
    if (sj * sj > floor2) return;
    const int first = complete_first(it, kept, floor2);
    double2* z = const_cast<double2*>(complete_col(it, j));
    // a deterministic pseudo-random start (the same for a matrix alone or in a batch)
    for (int r = threadIdx.x; r < it.p; r += 256) {
        const unsigned long long key = ((unsigned long long)j << 32) | (unsigned)r;
        z[r] = make_double2(splitmix_unit(2 * key), splitmix_unit(2 * key + 1));
    }
    __syncthreads();
    for (int pass = 0; pass < 2 && first > 0; ++pass) {
        complete_dots(it, z, 0, first, sDot);
        __syncthreads();
        complete_update(it, z, 0, first, sDot);
        __syncthreads();
    }
}

// per item: counts[2i] = first (the first frozen sorted position), counts[2i + 1] = kept
__global__ void k_complete_count(const CompleteItem* __restrict__ items, int count, int* __restrict__ counts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const CompleteItem it = items[i];
    const int kept = *it.kept;
    counts[2 * i] = complete_first(it, kept, it.floor_rel2 * (*it.fro2));
    counts[2 * i + 1] = kept;
}

// frozen positions with σ below 1e-280 get σ = 0 and Λ = 0 (their completed columns are
// stored as unit vectors; the others as unit vectors times σ)
__global__ void k_complete_zero_tiny(const CompleteItem* __restrict__ items, const int* __restrict__ counts) {
    const CompleteItem it = items[blockIdx.x];
    for (int j = counts[2 * blockIdx.x] + threadIdx.x; j < counts[2 * blockIdx.x + 1]; j += blockDim.x)
        if (!(it.sigma_sorted[j] >= 1e-280)) {
            it.sigma_sorted[j] = 0.0;
            it.lam_out[j] = 0.0;
        }
}

void launch_complete_project(const CompleteItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s) {
    if (ntiles > 0) k_complete_project<<<ntiles, 256, 0, s>>>(d_items, d_tiles);
}
void launch_complete_count(const CompleteItem* d_items, int count, int* d_counts, hipStream_t s) {
    if (count > 0) k_complete_count<<<(count + 255) / 256, 256, 0, s>>>(d_items, count, d_counts);
}
void launch_complete_zero_tiny(const CompleteItem* d_items, const int* d_counts, int count, hipStream_t s) {
    if (count > 0) k_complete_zero_tiny<<<count, 256, 0, s>>>(d_items, d_counts);
}

__global__ __launch_bounds__(256) void k_scatter_cols(const ScatterItem* __restrict__ items,
                                                      const int2* __restrict__ tiles) {
    const int2 tt = tiles[blockIdx.x];
    const ScatterItem it = items[tt.x];
    const int j = tt.y;
    if (j >= *it.kept) return;
    const double sg = it.sigma_sorted[j];
    const double f = (it.mul_sigma && sg > 0.0) ? sg : 1.0;
    double2* dst = it.x + (size_t)it.p * it.perm[j];
    const double2* src = it.f + (size_t)it.ldf * j;
    for (int r = threadIdx.x; r < it.p; r += 256) dst[r] = cscale(src[r], f);
}
void launch_scatter_cols(const ScatterItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s) {
    if (ntiles > 0) k_scatter_cols<<<ntiles, 256, 0, s>>>(d_items, d_tiles);
}

// ---------------------------------------------------------------------------
// write-back of kept columns (32 × 32 tiles: j = output column, c = row of src)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gather(const GatherItem* __restrict__ items, const int2* __restrict__ tiles) {
    __shared__ double2 sT[32][33];
    const int2 tt = tiles[blockIdx.x];
    const GatherItem g = items[tt.x];
    const int kept = *g.kept;
    const int ntc = (g.rows + 31) / 32;
    const int j0 = (tt.y / ntc) * 32, c0 = (tt.y % ntc) * 32;
    if (j0 >= kept) return;
    const int tid = threadIdx.x;
    const int cc = tid & 31;
    for (int s = 0; s < 4; ++s) {
        const int jj = (tid >> 5) + 8 * s;
        const int j = j0 + jj, c = c0 + cc;
        double2 v = cz();
        if (j < kept && c < g.rows) {
            const int src_col = g.perm ? g.perm[j] : j;
            v = g.src[(g.row_src ? g.row_src[c] : c) + (size_t)g.ld_src * src_col];
            double f = 1.0;
            if (g.divide_sigma) {
                const double sg = g.sigma_sorted[j];
                f = sg > 0.0 ? 1.0 / sg : 1.0;  // σ = 0: the column holds its unit vector
            }
            if (g.rowscale) f *= g.rowscale[(c / g.rowscale_div) % g.rowscale_mod];
            v = cscale(v, f);
        }
        if (!g.conj_transpose) {
            const size_t ldo = g.ld_out > 0 ? g.ld_out : g.rows;
            if (j < kept && c < g.rows) g.out[c + ldo * j] = v;
        } else {
            sT[jj][cc] = cconj(v);
        }
    }
    if (g.conj_transpose) {
        __syncthreads();
        for (int s = 0; s < 4; ++s) {
            const int c = c0 + (tid >> 5) + 8 * s;
            const int j = j0 + cc;
            if (j < kept && c < g.rows) g.out[j + (size_t)kept * c] = sT[cc][(tid >> 5) + 8 * s];
        }
    }
}
void launch_gather_items(const GatherItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s) {
    if (ntiles <= 0) return;
    k_gather<<<ntiles, 256, 0, s>>>(d_items, d_tiles);
}

}  // namespace qh
