// Device-side helpers: a minimal complex type over double2/float2 and the
// XCD-aware block-index mappings.
#pragma once

#include <hip/hip_runtime.h>

namespace hbx {

template <typename T> struct V2;
template <> struct V2<double> { using type = double2; };
template <> struct V2<float> { using type = float2; };
template <typename T> using c2 = typename V2<T>::type;

template <typename T> __device__ __forceinline__ c2<T> mk(T x, T y) { c2<T> r; r.x = x; r.y = y; return r; }
template <typename T> __device__ __forceinline__ c2<T> cadd(c2<T> a, c2<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ c2<T> csub(c2<T> a, c2<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ c2<T> cmul(c2<T> a, c2<T> b) {
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __device__ __forceinline__ c2<T> cconj(c2<T> a) { return mk<T>(a.x, -a.y); }
template <typename T> __device__ __forceinline__ c2<T> cscale(c2<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
// a * (-i)
template <typename T> __device__ __forceinline__ c2<T> cmul_mi(c2<T> a) { return mk<T>(a.y, -a.x); }
// a * (+i)
template <typename T> __device__ __forceinline__ c2<T> cmul_pi(c2<T> a) { return mk<T>(-a.y, a.x); }

__device__ __forceinline__ void sincospi_t(double a, double *s, double *c) { sincospi(a, s, c); }
__device__ __forceinline__ void sincospi_t(float a, float *s, float *c) { sincospif(a, s, c); }

// exp(i*pi*a) for a given as a rational q/d (exact reduction done by caller)
template <typename T> __device__ __forceinline__ c2<T> expipi(T a) {
    T s, c; sincospi_t(a, &s, &c); return mk<T>(c, s);
}

// Load through the constant address space: for memory no kernel running at
// the same time writes (plan tables, inputs written by an earlier launch), so
// a wave-uniform address becomes a scalar load even when the pointer itself
// was read from memory (a table of per-group pointers) rather than passed
// as a kernel argument.
template <typename X> __device__ __forceinline__ X ldc(const X *p) {
    return *(const __attribute__((address_space(4))) X *)(p);
}
// the same for a complex value, one component at a time
template <typename T> __device__ __forceinline__ c2<T> ldc2(const c2<T> *p) {
    const T *q = reinterpret_cast<const T *>(p);
    return mk<T>(ldc(q), ldc(q + 1));
}

// Loads and stores through the global address space, for pointers read from
// memory (a table of per-group pointers), which the compiler would otherwise
// treat as generic (flat) addresses held whole in vector registers.
#define HBX_GLOBAL __attribute__((address_space(1)))
template <typename X> __device__ __forceinline__ X ldg(const X *p) { return *(const HBX_GLOBAL X *)(p); }
template <typename T> __device__ __forceinline__ c2<T> ldg2(const c2<T> *p) {
    const T *q = reinterpret_cast<const T *>(p);
    return mk<T>(ldg(q), ldg(q + 1));
}
template <typename X> __device__ __forceinline__ void stg(X *p, X v) { *(HBX_GLOBAL X *)(p) = v; }
template <typename T> __device__ __forceinline__ void stg2(c2<T> *p, c2<T> v) {
    T *q = reinterpret_cast<T *>(p);
    stg(q, v.x);
    stg(q + 1, v.y);
}

// Wave-uniform copy of a value (lets the compiler keep it in an SGPR).
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// XCD-aware mapping of a 1-D block id onto (group, item) so that the `per`
// consecutive items of one group run on the same XCD back to back.  MI355X
// dispatches consecutive workgroups round-robin over its 8 XCDs.
//   g = 8*k + xcd ;  group = 8*(k / per) + xcd ;  item = k % per
__device__ __forceinline__ void xcd_map(int g, int per, int *group, int *item) {
    int xcd = g & 7, k = g >> 3;
    *group = 8 * (k / per) + xcd;
    *item = k % per;
}

}  // namespace hbx
