// Small kernels of the MPO path: exact site-wise application, a rectangular
// identity, and a complex dot product.
#include "qh_kernels.h"
#include "qh_complex.h"

#include <algorithm>

namespace qh {

// out[(b + chb w), o, (c + chc v)] = Σ_i W[w, o, i, v] A[b, i, c]   (layouts first index fastest)
__global__ void k_mpo_apply(const double2* __restrict__ w, const double2* __restrict__ a, double2* __restrict__ out,
                            int chb, int chc, int dl, int dr, int d) {
    const size_t lrows = (size_t)chb * dl;
    const size_t total = lrows * d * (size_t)chc * dr;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t l = e % lrows;
        const int o = (int)((e / lrows) % d);
        const size_t r = e / (lrows * d);
        const int b = (int)(l % chb), ww = (int)(l / chb);
        const int c = (int)(r % chc), v = (int)(r / chc);
        double2 acc = cz();
        for (int i = 0; i < d; ++i)
            acc = cfma(w[ww + (size_t)dl * (o + (size_t)d * (i + (size_t)d * v))],
                       a[b + (size_t)chb * (i + (size_t)d * c)], acc);
        out[e] = acc;
    }
}
void launch_mpo_apply(const double2* w, const double2* a, double2* out, int chb, int chc, int dl, int dr, int d,
                      hipStream_t s) {
    const size_t total = (size_t)chb * dl * d * chc * dr;
    if (total == 0) return;
    const size_t blocks = std::min<size_t>((total + 255) / 256, 65536);
    k_mpo_apply<<<(int)blocks, 256, 0, s>>>(w, a, out, chb, chc, dl, dr, d);
}

// p (m × n, ld m) <- [I; 0]
__global__ void k_eye_rect(double2* __restrict__ p, int m, int n) {
    const size_t total = (size_t)m * n;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t i = e % m, j = e / m;
        p[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    }
}
void launch_eye_rect(double2* p, int m, int n, hipStream_t s) {
    const size_t total = (size_t)m * n;
    if (total == 0) return;
    const size_t blocks = std::min<size_t>((total + 255) / 256, 65536);
    k_eye_rect<<<(int)blocks, 256, 0, s>>>(p, m, n);
}

// *out = Σ conj(x_k) y_k  (one workgroup, fixed order)
__global__ __launch_bounds__(256) void k_dot(const double2* __restrict__ x, const double2* __restrict__ y, size_t n,
                                             double2* __restrict__ out) {
    __shared__ double2 sRed[256];
    double2 s = cz();
    for (size_t e = threadIdx.x; e < n; e += 256) s = cfmac(x[e], y[e], s);
    sRed[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sRed[threadIdx.x] = cadd(sRed[threadIdx.x], sRed[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sRed[0];
}
void launch_dot(const double2* x, const double2* y, size_t n, double2* out, hipStream_t s) {
    k_dot<<<1, 256, 0, s>>>(x, y, n, out);
}

// *out = Σ_p e[p + n p]  (trace of an n × n matrix, one workgroup, fixed order)
__global__ __launch_bounds__(256) void k_trace(const double2* __restrict__ e, int n, double2* __restrict__ out) {
    __shared__ double2 sRed[256];
    double2 s = cz();
    for (int p = threadIdx.x; p < n; p += 256) s = cadd(s, e[(size_t)p * (n + 1)]);
    sRed[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sRed[threadIdx.x] = cadd(sRed[threadIdx.x], sRed[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sRed[0];
}
void launch_trace(const double2* e, int n, double2* out, hipStream_t s) { k_trace<<<1, 256, 0, s>>>(e, n, out); }

}  // namespace qh
