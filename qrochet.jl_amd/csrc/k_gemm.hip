// Two-site contraction (θ build with fused gate) and a batched complex GEMM,
// FP64 on the vector ALUs with LDS-staged tiles.
//
// θ[(l,a),(b,r)] = rs[l] cs[r] Σ_{o1,o2} G[a,b,o1,o2] Σ_m A_l[l,o1,m] λ[m] A_r[m,o2,r]
// (Qrochet `evolve_2site!`: contract!(tn, bond) then contract!(tn, gate inds),
//  src/Ansatz/Chain.jl:616,636, Vidal variant :669-685).  With the (l, o, r)
// device layout the product is a plain column-major GEMM (M = d·χl, K = χm,
// N = d·χr); a tile covers a range of l and r for *all* physical indices so the
// 4×4 gate mixes values a thread already holds.
#include "qh_kernels.h"
#include "qh_complex.h"

namespace qh {

// ---------------------------------------------------------------------------
// θ build
// ---------------------------------------------------------------------------
template <int D, int I>
__global__ __launch_bounds__(256) void k_theta_f64(const ThetaItem* __restrict__ items,
                                                   const int2* __restrict__ tiles) {
    constexpr int TL = 16 * I, TR = 16 * I, KC = 16;
    constexpr int BST = D * TR + 1;
    __shared__ double2 sA[KC][D * TL];
    __shared__ double2 sB[KC][BST];
    __shared__ double2 sG[D * D * D * D];

    const int tid = threadIdx.x;
    const int2 tt = tiles[blockIdx.x];
    const ThetaItem it = items[tt.x];
    const int l0 = (tt.y % it.tiles_l) * TL;
    const int r0 = (tt.y / it.tiles_l) * TR;
    const int chil = it.chil, chim = it.chim, chir = it.chir;
    const int lt = tid & 15, rt = tid >> 4;

    if (tid < D * D * D * D) sG[tid] = it.gate[tid];

    double2 acc[I][I][D][D];
#pragma unroll
    for (int i = 0; i < I; ++i)
#pragma unroll
        for (int j = 0; j < I; ++j)
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = 0; b < D; ++b) acc[i][j][a][b] = cz();

    for (int m0 = 0; m0 < chim; m0 += KC) {
        for (int e = tid; e < KC * D * TL; e += 256) {
            const int l = e % TL, o = (e / TL) % D, kk = e / (TL * D);
            const int gl = l0 + l, gm = m0 + kk;
            double2 v = cz();
            if (gl < chil && gm < chim) {
                v = it.al[gl + (size_t)chil * (o + D * (size_t)gm)];
                if (it.lam) v = cscale(v, it.lam[gm]);
            }
            sA[kk][o * TL + l] = v;
        }
        for (int e = tid; e < KC * D * TR; e += 256) {
            const int kk = e % KC, o = (e / KC) % D, r = e / (KC * D);
            const int gr = r0 + r, gm = m0 + kk;
            double2 v = cz();
            if (gr < chir && gm < chim) v = it.ar[gm + (size_t)chim * (o + D * (size_t)gr)];
            sB[kk][o * TR + r] = v;
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < KC; ++kk) {
            double2 av[I][D], bv[I][D];
#pragma unroll
            for (int i = 0; i < I; ++i)
#pragma unroll
                for (int o = 0; o < D; ++o) {
                    av[i][o] = sA[kk][o * TL + lt + 16 * i];
                    bv[i][o] = sB[kk][o * TR + rt + 16 * i];
                }
#pragma unroll
            for (int i = 0; i < I; ++i)
#pragma unroll
                for (int j = 0; j < I; ++j)
#pragma unroll
                    for (int a = 0; a < D; ++a)
#pragma unroll
                        for (int b = 0; b < D; ++b) acc[i][j][a][b] = cfma(av[i][a], bv[j][b], acc[i][j][a][b]);
        }
        __syncthreads();
    }

    const int m = D * chil, n = D * chir;
#pragma unroll
    for (int i = 0; i < I; ++i) {
        const int l = l0 + lt + 16 * i;
        if (l >= chil) continue;
        const double sl = it.rs ? it.rs[l] : 1.0;
#pragma unroll
        for (int j = 0; j < I; ++j) {
            const int r = r0 + rt + 16 * j;
            if (r >= chir) continue;
            const double sc = sl * (it.cs ? it.cs[r] : 1.0);
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = 0; b < D; ++b) {
                    double2 v = cz();
#pragma unroll
                    for (int o1 = 0; o1 < D; ++o1)
#pragma unroll
                        for (int o2 = 0; o2 < D; ++o2)
                            v = cfma(sG[a + D * b + D * D * (o1 + D * o2)], acc[i][j][o1][o2], v);
                    v = cscale(v, sc);
                    const size_t row = (size_t)l + (size_t)chil * a;
                    const size_t col = (size_t)b + (size_t)D * r;
                    if (!it.transposed)
                        it.x[row + (size_t)m * col] = v;
                    else
                        it.x[col + (size_t)n * row] = cconj(v);
                }
        }
    }
}

void launch_theta(const ThetaItem* d_items, const int2* d_tiles, int ntiles, int d, hipStream_t s) {
    if (ntiles <= 0) return;
    switch (d) {
        case 2: k_theta_f64<2, 2><<<ntiles, 256, 0, s>>>(d_items, d_tiles); break;
        case 3: k_theta_f64<3, 1><<<ntiles, 256, 0, s>>>(d_items, d_tiles); break;
        case 4: k_theta_f64<4, 1><<<ntiles, 256, 0, s>>>(d_items, d_tiles); break;
        default: break;
    }
}

int theta_tile_extent(int d) { return d == 2 ? 32 : 16; }

// ---------------------------------------------------------------------------
// batched GEMM  C = alpha op(A) diag(dk) op(B) (+ beta C), then row/col scales
// ---------------------------------------------------------------------------
// one 64 × 64 output tile of item g
__device__ __forceinline__ void gemm_tile_f64(const GemmItem& g, int tile) {
    constexpr int TM = GEMM_TM, TN = GEMM_TN, TK = GEMM_TK;
    __shared__ double2 sA[TK][TM + 1];
    __shared__ double2 sB[TK][TN + 1];
    const int tid = threadIdx.x;
    const int m0 = (tile % g.tiles_m) * TM, n0 = (tile / g.tiles_m) * TN;
    const int tm = tid & 15, tn = tid >> 4;

    double2 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = cz();

    for (int k0 = 0; k0 < g.k; k0 += TK) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int e = tid + 256 * s;
            int row, kk;
            if (op_plain(g.opa)) { row = e & 63; kk = e >> 6; }
            else { kk = e & 15; row = e >> 4; }
            const int gr = m0 + row, gk = k0 + kk;
            double2 v = cz();
            if (gr < g.m && gk < g.k) {
                v = load_op(g.a, g.lda, gr, gk, g.opa);
                if (g.dk) v = cscale(v, g.dk[((gk + g.dk_off) / g.dk_div) % g.dk_mod]);
            }
            sA[kk][row] = v;
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int e = tid + 256 * s;
            int col, kk;
            if (op_plain(g.opb)) { kk = e & 15; col = e >> 4; }
            else { col = e & 63; kk = e >> 6; }
            const int gc = n0 + col, gk = k0 + kk;
            double2 v = cz();
            if (gc < g.n && gk < g.k) v = load_op(g.b, g.ldb, gk, gc, g.opb);
            sB[kk][col] = v;
        }
        __syncthreads();
#pragma unroll 4
        for (int kk = 0; kk < TK; ++kk) {
            double2 av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) av[i] = sA[kk][tm + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) bv[j] = sB[kk][tn + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = cfma(av[i], bv[j], acc[i][j]);
        }
        __syncthreads();
    }
    const bool use_beta = (g.beta.x != 0.0 || g.beta.y != 0.0);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = m0 + tm + 16 * i;
        if (r >= g.m) continue;
        const double rsc = g.rowscale ? g.rowscale[r] : 1.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int c = n0 + tn + 16 * j;
            if (c >= g.n) continue;
            double2 v = cmul(g.alpha, acc[i][j]);
            v = cscale(v, rsc * (g.colscale ? g.colscale[c] : 1.0));
            double2* dst = g.c + r + (size_t)g.ldc * c;
            if (use_beta) v = cfma(g.beta, *dst, v);
            *dst = v;
        }
    }
}

__global__ __launch_bounds__(256) void k_gemm_reduce(const GemmReduceItem* __restrict__ items,
                                                     const int2* __restrict__ tiles) {
    const int2 tt = tiles[blockIdx.x];
    const GemmReduceItem g = items[tt.x];
    const size_t mn = (size_t)g.m * g.n;
    const size_t e = (size_t)tt.y * 256 + threadIdx.x;
    if (e >= mn) return;
    double2 v = g.w[e];
    for (int s = 1; s < g.slices; ++s) v = cadd(v, g.w[e + mn * s]);
    const int r = (int)(e % g.m), c = (int)(e / g.m);
    v = cmul(g.alpha, v);
    v = cscale(v, (g.rowscale ? g.rowscale[r] : 1.0) * (g.colscale ? g.colscale[c] : 1.0));
    double2* dst = g.c + r + (size_t)g.ldc * c;
    if (g.beta.x != 0.0 || g.beta.y != 0.0) v = cfma(g.beta, *dst, v);
    *dst = v;
}
void launch_gemm_reduce(const GemmReduceItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s) {
    if (ntiles > 0) k_gemm_reduce<<<ntiles, 256, 0, s>>>(d_items, d_tiles);
}

__global__ __launch_bounds__(256) void k_gemm_f64(const GemmItem* __restrict__ items, const int2* __restrict__ tiles) {
    const int2 tt = tiles[blockIdx.x];
    const GemmItem g = items[tt.x];
    gemm_tile_f64(g, tt.y);
}

// one GEMM whose descriptor travels as a kernel argument (no descriptor upload)
__global__ __launch_bounds__(256) void k_gemm_f64_arg(GemmItem g) { gemm_tile_f64(g, blockIdx.x); }

void launch_gemm(const GemmItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s) {
    if (ntiles <= 0) return;
    k_gemm_f64<<<ntiles, 256, 0, s>>>(d_items, d_tiles);
}

void launch_gemm_arg(GemmItem g, hipStream_t s) {
    g.tiles_m = (g.m + GEMM_TM - 1) / GEMM_TM;
    g.tiles_n = (g.n + GEMM_TN - 1) / GEMM_TN;
    if (g.dk_mod <= 0) g.dk_mod = 1;
    if (g.dk_div <= 0) g.dk_div = 1;
    const int ntiles = g.tiles_m * g.tiles_n;
    if (ntiles <= 0) return;
    k_gemm_f64_arg<<<ntiles, 256, 0, s>>>(g);
}

}  // namespace qh
