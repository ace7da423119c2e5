// Memory-bound helper kernels of the MPS update path: operator application on a
// physical index, Λ absorption (`contract!(…, :between)`, src/Ansatz/Chain.jl:309-333)
// and pinv(Λ) scaling (`unpack_2sitewf!` :708-713, `canonize!` :488-494),
// conjugate transposes, reductions for canonicality tests and the closing dot
// of the environment sweeps.
#include "qh_kernels.h"
#include "qh_complex.h"

namespace qh {

// T[l, μ, r] <- Σ_ν Op[μ + D ν] T[l, ν, r]   (D <= 16), one thread per (l, r)
template <int D>
__global__ __launch_bounds__(256) void k_apply_op(double2* __restrict__ t, int left, int right,
                                                  const double2* __restrict__ op) {
    __shared__ double2 sOp[D * D];
    for (int i = threadIdx.x; i < D * D; i += blockDim.x) sOp[i] = op[i];
    __syncthreads();
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)left * right;
    if (e >= total) return;
    const size_t l = e % left, r = e / left;
    double2* base = t + l + (size_t)left * D * r;
    double2 in[D];
#pragma unroll
    for (int v = 0; v < D; ++v) in[v] = base[(size_t)left * v];
#pragma unroll
    for (int u = 0; u < D; ++u) {
        double2 acc = cz();
#pragma unroll
        for (int v = 0; v < D; ++v) acc = cfma(sOp[u + D * v], in[v], acc);
        base[(size_t)left * u] = acc;
    }
}

void launch_apply_op(double2* t, int left, int dim, int right, const double2* op, hipStream_t s) {
    const size_t total = (size_t)left * right;
    if (total == 0) return;
    const int grid = (int)((total + 255) / 256);
    switch (dim) {
        case 2: k_apply_op<2><<<grid, 256, 0, s>>>(t, left, right, op); break;
        case 3: k_apply_op<3><<<grid, 256, 0, s>>>(t, left, right, op); break;
        case 4: k_apply_op<4><<<grid, 256, 0, s>>>(t, left, right, op); break;
        case 8: k_apply_op<8><<<grid, 256, 0, s>>>(t, left, right, op); break;
        case 9: k_apply_op<9><<<grid, 256, 0, s>>>(t, left, right, op); break;
        case 16: k_apply_op<16><<<grid, 256, 0, s>>>(t, left, right, op); break;
        default: break;
    }
}

// The same for a batch of tensors with their own shapes and operators (D <= 16, read
// at run time); tile map: (item, block of 256 (l, r) pairs).
__global__ __launch_bounds__(256) void k_apply_op_batch(const ApplyItem* __restrict__ items,
                                                        const int2* __restrict__ tiles) {
    __shared__ double2 sOp[256];
    const int2 tt = tiles[blockIdx.x];
    const ApplyItem it = items[tt.x];
    const int D = it.dim;
    for (int i = threadIdx.x; i < D * D; i += blockDim.x) sOp[i] = it.op[i];
    __syncthreads();
    const size_t e = (size_t)tt.y * 256 + threadIdx.x;
    const size_t total = (size_t)it.left * it.right;
    if (e >= total) return;
    const size_t l = e % it.left, r = e / it.left;
    double2* base = it.t + l + (size_t)it.left * D * r;
    double2 in[16];
    for (int v = 0; v < D; ++v) in[v] = base[(size_t)it.left * v];
    for (int u = 0; u < D; ++u) {
        double2 acc = cz();
        for (int v = 0; v < D; ++v) acc = cfma(sOp[u + D * v], in[v], acc);
        base[(size_t)it.left * u] = acc;
    }
}
void launch_apply_op_batch(const ApplyItem* d_items, const int2* d_tiles, int ntiles, hipStream_t s) {
    if (ntiles > 0) k_apply_op_batch<<<ntiles, 256, 0, s>>>(d_items, d_tiles);
}

// A[l, o, r] *= f(λ[l]) (mode 0) or f(λ[r]) (mode 2);  f = pinv ? (|λ|>atol ? 1/λ : 0) : λ
__global__ void k_scale_bond(double2* __restrict__ a, int chil, int d, int chir, const double* __restrict__ lam,
                             int mode, int pinv, double atol) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)chil * d * chir;
    if (e >= total) return;
    const int idx = mode == 0 ? (int)(e % chil) : (int)(e / ((size_t)chil * d));
    double f = lam[idx];
    if (pinv) {
        const double inv = 1.0 / f;
        f = (fabs(f) > atol && isfinite(inv)) ? inv : 0.0;
    }
    a[e] = cscale(a[e], f);
}
void launch_scale_bond(cplx* a, int chil, int d, int chir, const double* scale, int mode, int pinv, double atol,
                       hipStream_t s) {
    const size_t total = (size_t)chil * d * chir;
    if (total == 0) return;
    k_scale_bond<<<(int)((total + 255) / 256), 256, 0, s>>>(a, chil, d, chir, scale, mode, pinv, atol);
}

// out[i] = |λ_i| > atol ? 1/λ_i : 0   (`pinv(Diagonal(λ); atol)`)
__global__ void k_pinv(const double* __restrict__ lam, int n, double atol, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double f = lam[i];
    const double inv = 1.0 / f;
    out[i] = (fabs(f) > atol && isfinite(inv)) ? inv : 0.0;
}
void launch_pinv(const double* lam, int n, double atol, double* out, hipStream_t s) {
    if (n <= 0) return;
    k_pinv<<<(n + 255) / 256, 256, 0, s>>>(lam, n, atol, out);
}

// dst (cols × rows) = src^H,  src rows × cols (ld rows)
__global__ __launch_bounds__(256) void k_conj_transpose(const double2* __restrict__ src, int rows, int cols,
                                                        double2* __restrict__ dst) {
    __shared__ double2 sT[32][33];
    const int tr = blockIdx.x, tc = blockIdx.y;
    const int lane = threadIdx.x & 31, grp = threadIdx.x >> 5;
    for (int s = 0; s < 4; ++s) {
        const int r = tr * 32 + lane, c = tc * 32 + grp + 8 * s;
        if (r < rows && c < cols) sT[grp + 8 * s][lane] = cconj(src[r + (size_t)rows * c]);
    }
    __syncthreads();
    for (int s = 0; s < 4; ++s) {
        const int c = tc * 32 + lane, r = tr * 32 + grp + 8 * s;
        if (r < rows && c < cols) dst[c + (size_t)cols * r] = sT[lane][grp + 8 * s];
    }
}
void launch_conj_transpose(const cplx* src, int rows, int cols, cplx* dst, hipStream_t s) {
    if (rows <= 0 || cols <= 0) return;
    dim3 grid((rows + 31) / 32, (cols + 31) / 32);
    k_conj_transpose<<<grid, 256, 0, s>>>(src, rows, cols, dst);
}

// ||G - I||_F for an n × n matrix (single workgroup)
__global__ __launch_bounds__(1024) void k_gram_dev(const double2* __restrict__ g, int n, double* out) {
    __shared__ double sRed[16];
    double acc = 0.0;
    const size_t total = (size_t)n * n;
    for (size_t e = threadIdx.x; e < total; e += blockDim.x) {
        double2 v = g[e];
        if (e % n == e / n) v.x -= 1.0;
        acc += cabs2(v);
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) sRed[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sRed[w];
        *out = sqrt(t);
    }
}
void launch_gram_dev(const cplx* gram, int n, double* dev_out, hipStream_t s) {
    k_gram_dev<<<1, 1024, 0, s>>>(gram, n, dev_out);
}

// out = Σ_{b,a} E[b + nb a] R[a + na b] λb[b] λa[a]   (environment closing dot)
__global__ __launch_bounds__(1024) void k_env_close(const double2* __restrict__ e, const double2* __restrict__ r,
                                                    int nb, int na, const double* __restrict__ lb,
                                                    const double* __restrict__ la, double2* out) {
    __shared__ double sRe[16], sIm[16];
    double2 acc = cz();
    const size_t total = (size_t)nb * na;
    for (size_t k = threadIdx.x; k < total; k += blockDim.x) {
        const int b = (int)(k % nb), a = (int)(k / nb);
        double2 v = cmul(e[k], r[a + (size_t)na * b]);
        double f = 1.0;
        if (lb) f *= lb[b];
        if (la) f *= la[a];
        acc = cadd(acc, cscale(v, f));
    }
    for (int off = 32; off > 0; off >>= 1) {
        acc.x += __shfl_xor(acc.x, off, 64);
        acc.y += __shfl_xor(acc.y, off, 64);
    }
    if ((threadIdx.x & 63) == 0) { sRe[threadIdx.x >> 6] = acc.x; sIm[threadIdx.x >> 6] = acc.y; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double2 t = cz();
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { t.x += sRe[w]; t.y += sIm[w]; }
        *out = t;
    }
}
void launch_env_close(const cplx* e, const cplx* r, int nb, int na, const double* lb, const double* la, cplx* out,
                      hipStream_t s) {
    k_env_close<<<1, 1024, 0, s>>>(e, r, nb, na, lb, la, out);
}
// one workgroup per item, the same reduction order as k_env_close
__global__ __launch_bounds__(1024) void k_env_close_batch(const CloseItem* __restrict__ items) {
    const CloseItem it = items[blockIdx.x];
    __shared__ double sRe[16], sIm[16];
    double2 acc = cz();
    const size_t total = (size_t)it.nb * it.na;
    for (size_t k = threadIdx.x; k < total; k += blockDim.x) {
        const int b = (int)(k % it.nb), a = (int)(k / it.nb);
        double2 v = cmul(it.e[k], it.r[a + (size_t)it.na * b]);
        double f = 1.0;
        if (it.lb) f *= it.lb[b];
        if (it.la) f *= it.la[a];
        acc = cadd(acc, cscale(v, f));
    }
    for (int off = 32; off > 0; off >>= 1) {
        acc.x += __shfl_xor(acc.x, off, 64);
        acc.y += __shfl_xor(acc.y, off, 64);
    }
    if ((threadIdx.x & 63) == 0) { sRe[threadIdx.x >> 6] = acc.x; sIm[threadIdx.x >> 6] = acc.y; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double2 t = cz();
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) { t.x += sRe[w]; t.y += sIm[w]; }
        *it.out = t;
    }
}
void launch_env_close_batch(const CloseItem* d_items, int count, hipStream_t s) {
    if (count > 0) k_env_close_batch<<<count, 1024, 0, s>>>(d_items);
}

__global__ void k_fill(double* p, size_t n, double v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
void launch_fill(double* p, size_t n, double v, hipStream_t s) {
    if (n == 0) return;
    k_fill<<<(int)((n + 255) / 256), 256, 0, s>>>(p, n, v);
}

// Uniform [0, 1) entries for the sampler of `rand(Chain, Open, State)`
// (src/Ansatz/Chain.jl:223-256): entry e of buffer `key` is a splitmix64 hash of
// (seed, key, e), real part and -- for a complex element type -- imaginary part.
__device__ __forceinline__ double splitmix01(unsigned long long k) {
    k += 0x9e3779b97f4a7c15ull;
    k = (k ^ (k >> 30)) * 0xbf58476d1ce4e5b9ull;
    k = (k ^ (k >> 27)) * 0x94d049bb133111ebull;
    k ^= k >> 31;
    return (double)(k >> 11) * 0x1.0p-53;
}
__global__ void k_rand_uniform(double2* p, size_t n, unsigned long long seed, unsigned long long key, int complex_entries) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long base = seed ^ (key * 0xd1b54a32d192ed03ull);
    const double re = splitmix01(base + 2 * (unsigned long long)i);
    const double im = complex_entries ? splitmix01(base + 2 * (unsigned long long)i + 1) : 0.0;
    p[i] = make_double2(re, im);
}
void launch_rand_uniform(cplx* p, size_t n, unsigned long long seed, unsigned long long key, int complex_entries,
                         hipStream_t s) {
    if (n == 0) return;
    k_rand_uniform<<<(int)((n + 255) / 256), 256, 0, s>>>(p, n, seed, key, complex_entries);
}

__global__ void k_scale_real(double2* p, size_t n, double f) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_double2(p[i].x * f, p[i].y * f);
}
void launch_scale_real(cplx* p, size_t n, double f, hipStream_t s) {
    if (n == 0) return;
    k_scale_real<<<(int)((n + 255) / 256), 256, 0, s>>>(p, n, f);
}

// column j of a rows × cols matrix times 10^(-decades · j / (cols - 1))
__global__ void k_scale_cols_geometric(double2* p, int rows, int cols, double decades) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * cols) return;
    const int j = (int)(i / rows);
    const double f = cols > 1 ? exp10(-decades * j / (cols - 1)) : 1.0;
    p[i] = make_double2(p[i].x * f, p[i].y * f);
}
void launch_scale_cols_geometric(cplx* p, int rows, int cols, double decades, hipStream_t s) {
    const size_t t = (size_t)rows * cols;
    if (t == 0) return;
    k_scale_cols_geometric<<<(int)((t + 255) / 256), 256, 0, s>>>(p, rows, cols, decades);
}

// identity matrix (n × n, ld n)
__global__ void k_eye(double2* p, int n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * n) return;
    p[i] = make_double2((i % n) == (i / n) ? 1.0 : 0.0, 0.0);
}
void launch_eye(cplx* p, int n, hipStream_t s) {
    const size_t t = (size_t)n * n;
    if (t == 0) return;
    k_eye<<<(int)((t + 255) / 256), 256, 0, s>>>(p, n);
}

}  // namespace qh
