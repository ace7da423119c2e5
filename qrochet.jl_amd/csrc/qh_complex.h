// Complex FP64 helpers for device code (interleaved re/im in double2).
#pragma once
#include <hip/hip_runtime.h>

namespace qh {

__device__ __forceinline__ double2 cz() { return make_double2(0.0, 0.0); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
// c + a * b
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {
    c.x = fma(a.x, b.x, c.x);
    c.x = fma(-a.y, b.y, c.x);
    c.y = fma(a.x, b.y, c.y);
    c.y = fma(a.y, b.x, c.y);
    return c;
}
// c - a * b
__device__ __forceinline__ double2 cfms(double2 a, double2 b, double2 c) {
    c.x = fma(-a.x, b.x, c.x);
    c.x = fma(a.y, b.y, c.x);
    c.y = fma(-a.x, b.y, c.y);
    c.y = fma(-a.y, b.x, c.y);
    return c;
}
// c + conj(a) * b
__device__ __forceinline__ double2 cfmac(double2 a, double2 b, double2 c) {
    c.x = fma(a.x, b.x, c.x);
    c.x = fma(a.y, b.y, c.x);
    c.y = fma(a.x, b.y, c.y);
    c.y = fma(-a.y, b.x, c.y);
    return c;
}
__device__ __forceinline__ double cabs2(double2 a) { return fma(a.x, a.x, a.y * a.y); }

// Loads and stores through the global address space.  The matrices live in device
// memory but reach the kernels as generic pointers inside descriptors, which compile
// to flat accesses; a flat access is counted by the LDS counter as well, so every
// wait for an LDS read also waits for the loads still in flight.  Global accesses
// are counted by the vector-memory counter alone.
typedef const __attribute__((address_space(1))) double* gconst_f64;
typedef __attribute__((address_space(1))) double* gmut_f64;
__device__ __forceinline__ double2 gload(const double2* base, size_t idx) {
    const gconst_f64 g = (gconst_f64)base;  // address-space cast (generic to global)
    return make_double2(g[2 * idx], g[2 * idx + 1]);
}
// base[idx] where ok, else 0
__device__ __forceinline__ double2 ld_or_zero(const double2* base, size_t idx, bool ok) {
    return ok ? gload(base, idx) : make_double2(0.0, 0.0);
}
__device__ __forceinline__ void gstore(double2* base, size_t idx, double2 v) {
    const gmut_f64 g = (gmut_f64)base;
    g[2 * idx] = v.x;
    g[2 * idx + 1] = v.y;
}

}  // namespace qh
