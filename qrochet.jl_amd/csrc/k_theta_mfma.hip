// Two-site contraction on the matrix cores in reduced precision: θ (same
// definition and epilogue as k_theta_f64, src/Ansatz/Chain.jl:616,636 and the
// Vidal variant :669-685) with the χm contraction on MFMA
//   * FP32:   v_mfma_f32_16x16x4_f32, operands rounded from FP64 to FP32;
//   * BF16X3: each FP32 operand split as hi + lo (bfloat16 each) and
//             x·y ≈ xh·yh + xh·yl + xl·yh on v_mfma_f32_16x16x16_bf16 (the
//             "bf16x3" product of Henry, Tang & Heinecke, ARITH 2019),
// with FP32 accumulation inside each chunk of KC = 16 values of m and the chunk
// sums gathered in FP64 (for d <= 3).  With one FP32 accumulator over all χm, the
// accumulated rounding of the sum, not the rounding of the operands, set the error:
// on saturated χ = 256 TEBD layers the FP32 form's ⟨Z_i⟩ moved by ~2e-5 from FP64,
// against ~1e-7 for FP32 operands summed exactly.  The 4×4 gate mix, the Vidal
// scales and the output θ are FP64; the SVD that follows is FP64 throughout.
//
// Complex products use four real MFMAs: Re += ar·br + (-ai)·bi, Im += ar·bi + ai·br.
// A workgroup (4 waves) covers 32 values of l × 32 of r × all (o1, o2); wave w
// owns the 16×16 (l, r) block (w % 2, w / 2) for every (o1, o2), so the gate mix
// of an output element uses values the lane already holds.
#include "qh_kernels.h"
#include "qh_complex.h"
#include "qrochet_hip.h"

namespace qh {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short bf16x4 __attribute__((ext_vector_type(4)));

// FP32 -> bfloat16 bits, round to nearest even
__device__ __forceinline__ unsigned short bf16_bits(float x) {
    unsigned int u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_value(unsigned short b) { return __uint_as_float((unsigned int)b << 16); }
// x ≈ hi + lo; returns the two bit patterns
__device__ __forceinline__ void bf16_split(float x, short& hi, short& lo) {
    const unsigned short h = bf16_bits(x);
    hi = (short)h;
    lo = (short)bf16_bits(x - bf16_value(h));
}

template <int D, int P>
__global__ __launch_bounds__(256) void k_theta_mfma(const ThetaItem* __restrict__ items,
                                                    const int2* __restrict__ tiles) {
    constexpr int TL = 32, TR = 32, KC = 16;
    __shared__ float2 sA[KC][D][TL];  // (A_l λ)[l, o, m]   (m = chunk row)
    __shared__ float2 sB[KC][D][TR];  //  A_r[m, o, r]
    __shared__ double2 sG[D * D * D * D];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int2 tt = tiles[blockIdx.x];
    const ThetaItem it = items[tt.x];
    const int l0 = (tt.y % it.tiles_l) * TL;
    const int r0 = (tt.y / it.tiles_l) * TR;
    const int chil = it.chil, chim = it.chim, chir = it.chir;
    const int lb = wv & 1, rb = wv >> 1;
    const int li = lane & 15, kq = lane >> 4;

    if (tid < D * D * D * D) sG[tid] = it.gate[tid];

    f32x4 accR[D][D], accI[D][D];
    // chunk sums gathered in FP64 (d <= 3; at d = 4 the 128 extra registers would spill)
    constexpr bool ACC64 = D <= 3;
    double sumR[ACC64 ? D : 1][ACC64 ? D : 1][4], sumI[ACC64 ? D : 1][ACC64 ? D : 1][4];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) {
            accR[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
            accI[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
            if constexpr (ACC64)
#pragma unroll
                for (int t = 0; t < 4; ++t) sumR[a][b][t] = sumI[a][b][t] = 0.0;
        }

    for (int m0 = 0; m0 < chim; m0 += KC) {
        for (int e = tid; e < KC * D * TL; e += 256) {
            const int l = e % TL, o = (e / TL) % D, kk = e / (TL * D);
            const int gl = l0 + l, gm = m0 + kk;
            double2 v = cz();
            if (gl < chil && gm < chim) {
                v = it.al[gl + (size_t)chil * (o + D * (size_t)gm)];
                if (it.lam) v = cscale(v, it.lam[gm]);
            }
            sA[kk][o][l] = make_float2((float)v.x, (float)v.y);
        }
        for (int e = tid; e < KC * D * TR; e += 256) {
            const int kk = e % KC, o = (e / KC) % D, r = e / (KC * D);
            const int gr = r0 + r, gm = m0 + kk;
            double2 v = cz();
            if (gr < chir && gm < chim) v = it.ar[gm + (size_t)chim * (o + D * (size_t)gr)];
            sB[kk][o][r] = make_float2((float)v.x, (float)v.y);
        }
        __syncthreads();
        if constexpr (P == 1) {
            // lane (i, kq) holds A[i][k], B[k][j] for k = kq + 4 s
#pragma unroll
            for (int s = 0; s < KC / 4; ++s) {
                const int kk = kq + 4 * s;
                float2 av[D], bv[D];
#pragma unroll
                for (int o = 0; o < D; ++o) {
                    av[o] = sA[kk][o][16 * lb + li];
                    bv[o] = sB[kk][o][16 * rb + li];
                }
#pragma unroll
                for (int a = 0; a < D; ++a)
#pragma unroll
                    for (int b = 0; b < D; ++b) {
                        accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a].x, bv[b].x, accR[a][b], 0, 0, 0);
                        accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(-av[a].y, bv[b].y, accR[a][b], 0, 0, 0);
                        accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a].x, bv[b].y, accI[a][b], 0, 0, 0);
                        accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a].y, bv[b].x, accI[a][b], 0, 0, 0);
                    }
            }
        } else {
            // lane (i, kq) holds A[i][4 kq + t], B[4 kq + t][j], t = 0..3 (the whole chunk)
            bf16x4 arh[D], arl[D], aih[D], ail[D], ainh[D], ainl[D];
            bf16x4 brh[D], brl[D], bih[D], bil[D];
#pragma unroll
            for (int o = 0; o < D; ++o)
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const float2 a = sA[4 * kq + t][o][16 * lb + li];
                    const float2 b = sB[4 * kq + t][o][16 * rb + li];
                    short h, l;
                    bf16_split(a.x, h, l); arh[o][t] = h; arl[o][t] = l;
                    bf16_split(a.y, h, l); aih[o][t] = h; ail[o][t] = l;
                    ainh[o][t] = (short)(h ^ (short)0x8000); ainl[o][t] = (short)(l ^ (short)0x8000);
                    bf16_split(b.x, h, l); brh[o][t] = h; brl[o][t] = l;
                    bf16_split(b.y, h, l); bih[o][t] = h; bil[o][t] = l;
                }
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = 0; b < D; ++b) {
                    f32x4 r = accR[a][b], i = accI[a][b];
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh[a], brh[b], r, 0, 0, 0);
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh[a], brl[b], r, 0, 0, 0);
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arl[a], brh[b], r, 0, 0, 0);
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ainh[a], bih[b], r, 0, 0, 0);
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ainh[a], bil[b], r, 0, 0, 0);
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ainl[a], bih[b], r, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh[a], bih[b], i, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh[a], bil[b], i, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arl[a], bih[b], i, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aih[a], brh[b], i, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aih[a], brl[b], i, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ail[a], brh[b], i, 0, 0, 0);
                    accR[a][b] = r;
                    accI[a][b] = i;
                }
        }
        if constexpr (ACC64) {
#pragma unroll
            for (int a = 0; a < D; ++a)
#pragma unroll
                for (int b = 0; b < D; ++b) {
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        sumR[a][b][t] += (double)accR[a][b][t];
                        sumI[a][b][t] += (double)accI[a][b][t];
                    }
                    accR[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
                    accI[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
                }
        }
        __syncthreads();
    }

    // lane holds D[4 kq + t][li] of each 16×16 block: l = l0 + 16 lb + 4 kq + t, r = r0 + 16 rb + li
    const int m = D * chil, n = D * chir;
    const int r = r0 + 16 * rb + li;
    if (r >= chir) return;
    const double scr = it.cs ? it.cs[r] : 1.0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int l = l0 + 16 * lb + 4 * kq + t;
        if (l >= chil) continue;
        const double sc = scr * (it.rs ? it.rs[l] : 1.0);
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
            for (int b = 0; b < D; ++b) {
                double2 v = cz();
#pragma unroll
                for (int o1 = 0; o1 < D; ++o1)
#pragma unroll
                    for (int o2 = 0; o2 < D; ++o2)
                    {
                        const double2 acc = ACC64 ? make_double2(sumR[ACC64 ? o1 : 0][ACC64 ? o2 : 0][t],
                                                                 sumI[ACC64 ? o1 : 0][ACC64 ? o2 : 0][t])
                                                  : make_double2((double)accR[o1][o2][t], (double)accI[o1][o2][t]);
                        v = cfma(sG[a + D * b + D * D * (o1 + D * o2)], acc, v);
                    }
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

void launch_theta_mfma(const ThetaItem* d_items, const int2* d_tiles, int ntiles, int d, int precision,
                       hipStream_t s) {
    if (ntiles <= 0) return;
    const bool fp32 = precision == QH_PREC_FP32;
    switch (d) {
        case 2:
            if (fp32) k_theta_mfma<2, 1><<<ntiles, 256, 0, s>>>(d_items, d_tiles);
            else k_theta_mfma<2, 2><<<ntiles, 256, 0, s>>>(d_items, d_tiles);
            break;
        case 3:
            if (fp32) k_theta_mfma<3, 1><<<ntiles, 256, 0, s>>>(d_items, d_tiles);
            else k_theta_mfma<3, 2><<<ntiles, 256, 0, s>>>(d_items, d_tiles);
            break;
        case 4:
            if (fp32) k_theta_mfma<4, 1><<<ntiles, 256, 0, s>>>(d_items, d_tiles);
            else k_theta_mfma<4, 2><<<ntiles, 256, 0, s>>>(d_items, d_tiles);
            break;
        default: break;
    }
}

// ---------------------------------------------------------------------------
// Generic batched complex GEMM on MFMA (same items and tiles as k_gemm_f64):
// C = alpha · op(A) diag(dk) op(B) (+ beta C), rows × rowscale, columns × colscale.
// A 64×64 tile per workgroup; wave w computes the 32×32 quarter (w % 2, w / 2)
// as 2×2 MFMA blocks of 16×16.  Operands are staged through LDS in FP32; each
// k-chunk is accumulated in FP32 on the matrix cores and the chunk sums in FP64.
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(256) void k_gemm_mfma(const GemmItem* __restrict__ items,
                                                   const int2* __restrict__ tiles) {
    constexpr int TM = GEMM_TM, TN = GEMM_TN, TK = GEMM_TK;
    __shared__ float2 sA[TK][TM];
    __shared__ float2 sB[TK][TN];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int2 tt = tiles[blockIdx.x];
    const GemmItem g = items[tt.x];
    const int m0 = (tt.y % g.tiles_m) * TM, n0 = (tt.y / g.tiles_m) * TN;
    const int wm = (wv & 1) * 32, wn = (wv >> 1) * 32;
    const int li = lane & 15, kq = lane >> 4;

    f32x4 accR[2][2], accI[2][2];
    double sumR[2][2][4], sumI[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            accR[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
            accI[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int t = 0; t < 4; ++t) sumR[a][b][t] = sumI[a][b][t] = 0.0;
        }

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
            sA[kk][row] = make_float2((float)v.x, (float)v.y);
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
            sB[kk][col] = make_float2((float)v.x, (float)v.y);
        }
        __syncthreads();
        if constexpr (P == 1) {
#pragma unroll
            for (int st = 0; st < TK / 4; ++st) {
                const int kk = kq + 4 * st;
                float2 av[2], bv[2];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    av[t] = sA[kk][wm + 16 * t + li];
                    bv[t] = sB[kk][wn + 16 * t + li];
                }
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b) {
                        accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a].x, bv[b].x, accR[a][b], 0, 0, 0);
                        accR[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(-av[a].y, bv[b].y, accR[a][b], 0, 0, 0);
                        accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a].x, bv[b].y, accI[a][b], 0, 0, 0);
                        accI[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a].y, bv[b].x, accI[a][b], 0, 0, 0);
                    }
            }
        } else {
            bf16x4 arh[2], arl[2], aih[2], ail[2], ainh[2], ainl[2];
            bf16x4 brh[2], brl[2], bih[2], bil[2];
#pragma unroll
            for (int t2 = 0; t2 < 2; ++t2)
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const float2 a = sA[4 * kq + t][wm + 16 * t2 + li];
                    const float2 b = sB[4 * kq + t][wn + 16 * t2 + li];
                    short h, l;
                    bf16_split(a.x, h, l); arh[t2][t] = h; arl[t2][t] = l;
                    bf16_split(a.y, h, l); aih[t2][t] = h; ail[t2][t] = l;
                    ainh[t2][t] = (short)(h ^ (short)0x8000); ainl[t2][t] = (short)(l ^ (short)0x8000);
                    bf16_split(b.x, h, l); brh[t2][t] = h; brl[t2][t] = l;
                    bf16_split(b.y, h, l); bih[t2][t] = h; bil[t2][t] = l;
                }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    f32x4 r = accR[a][b], i = accI[a][b];
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh[a], brh[b], r, 0, 0, 0);
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh[a], brl[b], r, 0, 0, 0);
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arl[a], brh[b], r, 0, 0, 0);
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ainh[a], bih[b], r, 0, 0, 0);
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ainh[a], bil[b], r, 0, 0, 0);
                    r = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ainl[a], bih[b], r, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh[a], bih[b], i, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh[a], bil[b], i, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arl[a], bih[b], i, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aih[a], brh[b], i, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aih[a], brl[b], i, 0, 0, 0);
                    i = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ail[a], brh[b], i, 0, 0, 0);
                    accR[a][b] = r;
                    accI[a][b] = i;
                }
        }
        {  // chunk sums gathered in FP64
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        sumR[a][b][t] += (double)accR[a][b][t];
                        sumI[a][b][t] += (double)accI[a][b][t];
                    }
                    accR[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
                    accI[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
                }
        }
        __syncthreads();
    }
    const bool use_beta = (g.beta.x != 0.0 || g.beta.y != 0.0);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int c = n0 + wn + 16 * b + li;
            if (c >= g.n) continue;
            const double csc = g.colscale ? g.colscale[c] : 1.0;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int r = m0 + wm + 16 * a + 4 * kq + t;
                if (r >= g.m) continue;
                double2 v = cmul(g.alpha, make_double2(sumR[a][b][t], sumI[a][b][t]));
                v = cscale(v, csc * (g.rowscale ? g.rowscale[r] : 1.0));
                double2* dst = g.c + r + (size_t)g.ldc * c;
                if (use_beta) v = cfma(g.beta, *dst, v);
                *dst = v;
            }
        }
}

void launch_gemm_mfma(const GemmItem* d_items, const int2* d_tiles, int ntiles, int precision, hipStream_t s) {
    if (ntiles <= 0) return;
    if (precision == QH_PREC_FP32) k_gemm_mfma<1><<<ntiles, 256, 0, s>>>(d_items, d_tiles);
    else k_gemm_mfma<2><<<ntiles, 256, 0, s>>>(d_items, d_tiles);
}

}  // namespace qh
