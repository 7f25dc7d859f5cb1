// VALU Legendre wave bodies shared by the per-plan kernels (legendre.hip)
// and the persistent mixed-geometry kernels (multi.hip): one FP64 lambda
// recursion per ring pair and lane, contracted with the ring coefficients
// (analysis) or the scaled alm rows (synthesis).  Both callers run the same
// code, so their results agree bit for bit.
#pragma once

#include "hbx_device.h"
#include "hbx_internal.h"

namespace hbx {

constexpr int NEVER = 0x7fffffff;

// ---------------------------------------------------------------------------
// wave helpers (wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return uniform(v);
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return uniform(v);
}

// ---------------------------------------------------------------------------
// analysis: phase -> alm.  One wave per (m, ring-pair group, map); each lane
// holds R ring pairs ("slots", 64 consecutive pairs per slot), so a wave
// covers 64*R pairs and runs with no workgroup barrier.  Per l-chunk the 16
// partial sums (8 l x re/im) are reduced over the wave by an LDS transpose:
// every lane writes its 16 values, lane j then sums value j/4 over 16
// source lanes and two xor-shuffles finish the sum.  With a single group the
// wave writes s_lm * sum straight into alm; otherwise it writes raw partial
// sums per group and group_reduce_kernel adds them in a fixed order.
// ---------------------------------------------------------------------------
// transpose buffer of the VALU analysis: row v (one of the 2 CHUNK sums) of
// the 64 lanes.  The reduction reads 16 consecutive columns per lane, 4 lanes
// per row.  4-byte elements: rows 65 elements apart put those reads on
// distinct banks.  8-byte elements (two banks each): the 4 lanes of one row
// would meet on the same banks, so every 16 columns are followed by 4 pad
// elements and rows are 81 elements apart.
template <typename T> struct RedLayout;
template <> struct RedLayout<float> {
    static constexpr int ROW = 65;
    __device__ static int at(int v, int j) { return v * ROW + j; }
};
template <> struct RedLayout<double> {
    static constexpr int ROW = 81;
    __device__ static int at(int v, int j) { return v * ROW + j + 4 * (j >> 4); }
};
template <typename T> constexpr int red_elems() { return 2 * CHUNK * RedLayout<T>::ROW; }

// One wave's share of the VALU analysis: ring pairs grp*64R .. grp*64R+64R-1
// of one m and one map.  ph: the map's phase at column m of its m-block;
// dst: the map's alm (direct) or its partial-sum row of ring-pair group grp.
// red: red_elems<T>() elements of LDS owned by this wave.
template <typename T, int R>
__device__ __forceinline__ void
ana_wave(const c2<T> *__restrict__ ph, c2<T> *__restrict__ dst, bool direct, const int32_t *__restrict__ pair_n,
         const int32_t *__restrict__ pair_s, const double *__restrict__ px, const int64_t *__restrict__ rofs,
         const double *__restrict__ alpha, const double *__restrict__ sl, const int32_t *__restrict__ lseed,
         const double2 *__restrict__ seed, int npairs, int lmax, int m, int grp, int adjoint, T *red, int lane) {
    double x[R], u1[R], u2[R];
    c2<T> E[R], O[R];
    int ls[R], sst[R], sinj[R];
    const int p0 = grp * 64 * R + lane;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int p = p0 + 64 * k;
        u1[k] = 0.0;
        u2[k] = 0.0;
        if (p < npairs) {
            x[k] = ldg(px + p);
            ls[k] = ldg(lseed + (int64_t)m * npairs + p);
            const int rn = ldg(pair_n + p), rs = ldg(pair_s + p);
            const c2<T> fn = ldg2<T>(ph + (int64_t)rn * MB);
            const c2<T> fs = (rs != rn) ? ldg2<T>(ph + (int64_t)rs * MB) : mk<T>(0, 0);
            E[k] = cadd<T>(fn, fs);
            O[k] = csub<T>(fn, fs);
        } else {
            x[k] = 0.0;
            ls[k] = NEVER;
            E[k] = mk<T>(0, 0);
            O[k] = mk<T>(0, 0);
        }
        sst[k] = wave_min(ls[k]);
        sinj[k] = wave_max(ls[k] <= lmax ? ls[k] : -1);
    }
    constexpr int GS = R < 4 ? R : 4;
    int gst[R / GS];
    int lstart = NEVER;
#pragma unroll
    for (int k0 = 0; k0 < R; k0 += GS) {
        int v = NEVER;
#pragma unroll
        for (int k = k0; k < k0 + GS; ++k) v = min(v, sst[k]);
        gst[k0 / GS] = v;
        lstart = min(lstart, v);
    }

    const int64_t i0 = alm_index(m, m, lmax);
    const int64_t ro = ldc(rofs + m);
    // rows below every slot's switch-on are zero
    for (int l = m + lane; l < (lstart > lmax ? lmax + 1 : lstart); l += 64) stg2<T>(dst + i0 + (l - m), mk<T>(0, 0));

    const int v_id = lane >> 2, q0 = (lane & 3) * 16;
    for (int l0 = lstart; l0 <= lmax; l0 += CHUNK) {
        double a[CHUNK];
#pragma unroll
        for (int i = 0; i < CHUNK; ++i) a[i] = ldc(alpha + ro + (l0 - m) + i);
        T acc[2 * CHUNK];
#pragma unroll
        for (int i = 0; i < 2 * CHUNK; ++i) acc[i] = (T)0;
        // slots advance in groups of GS with their steps interleaved, so
        // GS independent recursion chains hide the FP64 latency
#pragma unroll
        for (int k0 = 0; k0 < R; k0 += GS) {
            if (l0 < gst[k0 / GS]) continue;
#pragma unroll
            for (int k = k0; k < k0 + GS; ++k) {
                if (l0 <= sinj[k] && ls[k] == l0) {
                    const double2 sd = ldg2<double>(seed + (int64_t)m * npairs + p0 + 64 * k);
                    u2[k] = sd.x;
                    u1[k] = sd.y;
                }
            }
#pragma unroll
            for (int i = 0; i < CHUNK; ++i) {
#pragma unroll
                for (int k = k0; k < k0 + GS; ++k) {
                    const double u = fma(a[i] * x[k], u1[k], -u2[k]);
                    u2[k] = u1[k];
                    u1[k] = u;
                    const T uu = (T)u;
                    const c2<T> d = (i & 1) ? O[k] : E[k];
                    acc[2 * i] = fma(uu, d.x, acc[2 * i]);
                    acc[2 * i + 1] = fma(uu, d.y, acc[2 * i + 1]);
                }
            }
        }
        // wave-level transpose-reduce of the 16 values
#pragma unroll
        for (int v = 0; v < 2 * CHUNK; ++v) red[RedLayout<T>::at(v, lane)] = acc[v];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        T s = (T)0;
#pragma unroll
        for (int q = 0; q < 16; ++q) s += red[RedLayout<T>::at(v_id, q0 + q)];
        s += __shfl_xor(s, 1, 64);
        s += __shfl_xor(s, 2, 64);
        if ((lane & 3) == 0) {
            const int i = v_id >> 1, c = v_id & 1;
            const int l = l0 + i;
            if (l <= lmax) {
                T val = s;
                if (direct) {
                    double dv = (double)s * sl[ro + (l0 - m) + i];
                    if (adjoint) {
                        if (m > 0) dv *= 2.0;
                        else if (c == 1) dv = 0.0;
                    }
                    val = (T)dv;
                }
                stg(reinterpret_cast<T *>(dst + i0 + (l - m)) + c, val);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// One wave's share of the VALU synthesis: ring pair p = p0 + lane of one m
// for NB maps that share each FP64 recursion step.  ar[j]: scaled row m of
// map j, indexed by l; out[j]: map j's synthesis phase at (m, half 0, p).
// Every map's sums are formed in the same order as with NB = 1.
template <typename T, int NB>
__device__ __forceinline__ void
syn_wave(const c2<T> *const *ar, c2<T> *const *out, const int32_t *__restrict__ pair_n,
         const int32_t *__restrict__ pair_s, const double *__restrict__ px, const double *__restrict__ al,
         const int32_t *__restrict__ lseed, const double2 *__restrict__ seed, int npairs, int lmax, int m, int p) {
    const bool valid = p < npairs;
    const double x = valid ? ldg(px + p) : 0.0;
    const int ls = valid ? ldg(lseed + (int64_t)m * npairs + p) : NEVER;
    const int lstart = wave_min(ls);
    const int linj = wave_max(ls <= lmax ? ls : -1);
    c2<T> ge[NB], go[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        ge[j] = mk<T>(0, 0);
        go[j] = mk<T>(0, 0);
    }
    if (lstart <= lmax) {
        double u2 = 0.0, u1 = 0.0;
        for (int l0 = lstart; l0 <= lmax; l0 += CHUNK) {
            if (l0 <= linj && ls == l0) {
                const double2 sd = ldg2<double>(seed + (int64_t)m * npairs + p);
                u2 = sd.x;
                u1 = sd.y;
            }
#pragma unroll
            for (int i = 0; i < CHUNK; ++i) {
                const double u = fma(ldc(al + l0 + i) * x, u1, -u2);
                u2 = u1;
                u1 = u;
                const T uu = (T)u;
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    const c2<T> av = ldc2<T>(ar[j] + l0 + i);
                    if (i & 1) {
                        go[j].x = fma(uu, av.x, go[j].x);
                        go[j].y = fma(uu, av.y, go[j].y);
                    } else {
                        ge[j].x = fma(uu, av.x, ge[j].x);
                        ge[j].y = fma(uu, av.y, ge[j].y);
                    }
                }
            }
        }
    }
    if (!valid) return;
    // synthesis layout: ring pair p, north (ring pair_n[p]) and south (pair_s[p]) halves
    const bool two = ldg(pair_s + p) != ldg(pair_n + p);
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        stg2<T>(out[j], cadd<T>(ge[j], go[j]));
        if (two) stg2<T>(out[j] + (int64_t)pairs8(npairs) * MB, csub<T>(ge[j], go[j]));
    }
}

// alm[idx(l,m)] = s_lm * sum_g partial[g][idx(l,m)] for one row m (fixed
// summation order), by the threads of one workgroup.  src: group 0's partial
// row m (groups nlm apart); dst: alm row m; slr: s_lm of row m indexed by
// l - m; with lfirst (per group of row m), group g's partial is read only for
// l >= lfirst[g] (zero below).
template <typename T>
__device__ __forceinline__ void reduce_row(const c2<T> *__restrict__ src, int64_t nlm, int ngroups,
                                           c2<T> *__restrict__ dst, const double *__restrict__ slr, int lmax, int m,
                                           int adjoint, const int32_t *__restrict__ lfirst) {
    for (int j = threadIdx.x; j <= lmax - m; j += blockDim.x) {
        double re = 0.0, im = 0.0;
        for (int g = 0; g < ngroups; ++g) {
            if (lfirst && m + j < lfirst[g]) continue;
            const c2<T> v = src[(int64_t)g * nlm + j];
            re += (double)v.x;
            im += (double)v.y;
        }
        double s = slr[j];
        if (adjoint) {
            if (m > 0) s *= 2.0;
            else im = 0.0;
        }
        dst[j] = mk<T>((T)(re * s), (T)(im * s));
    }
}

// synthesis pre-pass for one row m: scaled[l-m] = s_lm * a_lm (halved for
// m > 0 in the map2alm adjoint) over the row's len padded entries (zero past
// lmax, so recursion chunks never read past a row), by one workgroup.
template <typename T>
__device__ __forceinline__ void prescale_row(const c2<T> *__restrict__ a, c2<T> *__restrict__ d,
                                             const double *__restrict__ slr, int len, int lmax, int m, int adjoint) {
    const double f = (adjoint && m > 0) ? 0.5 : 1.0;
    for (int j = threadIdx.x; j < len; j += blockDim.x) {
        c2<T> v = mk<T>(0, 0);
        if (m + j <= lmax) v = cscale<T>(a[j], (T)(slr[j] * f));
        d[j] = v;
    }
}

}  // namespace hbx
