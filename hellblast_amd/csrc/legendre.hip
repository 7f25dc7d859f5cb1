// Legendre stage: lambda_lm(cos theta) recursion fused with the ring
// contraction (analysis) or the l-sum (synthesis).
//
// Reference: plm_recurrence (legendre.cu:2-35) writes the unnormalised P_lm
// for every (ring, l) of one m to HBM, the host multiplies by _alm_norm
// (sht_raw.py:26-28, overflows for l+m >= 171) and reduces with cp.sum over
// rings (sht_raw.py:53) or over l (isht.py:21).
//
// Here lambda_lm = s_lm * mu_lm with the rescaled three-term recursion
//     mu_l = alpha_lm * x * mu_{l-1} - mu_{l-2}            (2 VALU ops)
// where  a_lm = sqrt((4l^2-1)/(l^2-m^2)),  s_lm = (a_lm/a_{l-1,m}) s_{l-2,m},
// alpha_lm = a_lm s_{l-1,m}/s_lm  (s_mm = s_{m+1,m} = 1).  This equals
// lambda_l = a_l (x lambda_{l-1} - lambda_{l-2}/a_{l-1}) in exact arithmetic.
// lambda is never stored.  sin^m theta underflow is handled once per plan: a
// seed kernel runs the recursion in mantissa*2^E form and records, per
// (m, ring pair), the first chunk-aligned l at which |lambda| can exceed
// 2^tol ("switch-on") together with (mu_{l-2}, mu_{l-1}) there.  Transforms
// start each (m, ring pair) at its seed and never rescale.
//
// North/south mirror rings share x up to sign: lambda_lm(-x) =
// (-1)^{l-m} lambda_lm(x), so one recursion serves both rings:
//   analysis   E = F_n + F_s (l-m even), O = F_n - F_s (l-m odd)
//   synthesis  G_n = G_even + G_odd,  G_s = G_even - G_odd.
#include <algorithm>
#include <cstdlib>

#include "hbx_device.h"
#include "hbx_internal.h"

namespace hbx {

constexpr int NEVER = 0x7fffffff;

// ---------------------------------------------------------------------------
// plan-time kernels
// ---------------------------------------------------------------------------
// one thread per m: alpha_lm and s_lm for l = m .. padded row end
__global__ void recursion_tables_kernel(const int64_t *__restrict__ rofs, double *alpha, double *sl,
                                        int lmax, int mmax) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m > mmax) return;
    const int64_t o = rofs[m];
    const int len = (int)(rofs[m + 1] - o);
    double s2 = 1.0, s1 = 1.0, aprev = 0.0;   // s_{l-2}, s_{l-1}, a_{l-1}
    for (int j = 0; j < len; ++j) {
        const int l = m + j;
        double al = 0.0, s = 0.0;
        if (l > lmax) {
            al = 0.0; s = 0.0;
        } else if (j == 0) {
            al = 0.0; s = 1.0;
        } else {
            const double a = sqrt((4.0 * (double)l * l - 1.0) / ((double)(l - m) * (double)(l + m)));
            if (j == 1) {
                s = 1.0;
                al = a;
            } else {
                s = (a / aprev) * s2;
                al = a * s1 / s;
            }
            aprev = a;
            s2 = s1;
            s1 = s;
        }
        alpha[o + j] = al;
        sl[o + j] = s;
    }
}

// one thread per ring pair: lambda_mm = mant * 2^ex for m = 0..mmax
__global__ void lambda_mm_kernel(const double *__restrict__ sinp, double *mant, int *ex, int npairs,
                                 int mmax) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    const double st = sinp[p];
    int e = 0;
    double v = frexp(0.28209479177387814347403972578 /* 1/sqrt(4 pi) */, &e);
    mant[p] = v;
    ex[p] = e;
    for (int m = 1; m <= mmax; ++m) {
        v = v * (-sqrt((2.0 * m + 1.0) / (2.0 * m))) * st;
        int de = 0;
        v = frexp(v, &de);
        e += de;
        mant[(int64_t)m * npairs + p] = v;
        ex[(int64_t)m * npairs + p] = e;
    }
}

// one thread per (m, ring pair): switch-on seed of the recursion
__global__ void seed_kernel(const double *__restrict__ px, const double *__restrict__ mant,
                            const int *__restrict__ ex, const int64_t *__restrict__ rofs,
                            const double *__restrict__ alpha, const double *__restrict__ sl,
                            int32_t *lseed, double2 *seed, int npairs, int lmax, int tol_log2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    if (p >= npairs) return;
    const int64_t q = (int64_t)m * npairs + p;
    const double x = px[p];
    const double m0 = mant[q];
    const int e0 = ex[q];
    const double thr = ldexp(1.0, tol_log2);
    const double big = ldexp(1.0, 200), small = ldexp(1.0, -200);
    const double *al = alpha + rofs[m];
    const double *sv = sl + rofs[m];
    int lon = NEVER;
    if (m0 != 0.0) {
        if (fabs(ldexp(m0, e0)) >= thr) {
            lon = m;
        } else {
            double u2 = 0.0, u1 = m0;
            int e = e0;
            for (int l = m + 1; l <= lmax; ++l) {
                const double u = al[l - m] * x * u1 - u2;
                u2 = u1;
                u1 = u;
                if (fabs(u1) > big) { u1 *= small; u2 *= small; e += 200; }
                if (e > tol_log2 - 1100 && fabs(ldexp(sv[l - m] * u1, e)) >= thr) { lon = l; break; }
            }
        }
    }
    if (lon == NEVER) {
        lseed[q] = NEVER;
        seed[q] = make_double2(0.0, 0.0);
        return;
    }
    const int ls = m + ((lon - m) / CHUNK) * CHUNK;
    lseed[q] = ls;
    if (ls == m) {
        seed[q] = make_double2(-ldexp(m0, e0), 0.0);
        return;
    }
    double u2 = 0.0, u1 = m0;
    int e = e0;
    for (int l = m + 1; l <= ls - 1; ++l) {
        const double u = al[l - m] * x * u1 - u2;
        u2 = u1;
        u1 = u;
        if (fabs(u1) > big) { u1 *= small; u2 *= small; e += 200; }
    }
    seed[q] = make_double2(ldexp(u2, e), ldexp(u1, e));
}

int plan_build_recursion(Plan *p, const std::vector<double> &pair_cos, const std::vector<double> &pair_sin,
                         double tol_log2) {
    const int mmax = p->mmax, lmax = p->lmax, np = p->npairs;
    std::vector<int64_t> rofs(mmax + 2);
    rofs[0] = 0;
    auto rup = [](int v) { return (v + CHUNK - 1) / CHUNK * CHUNK; };
    for (int m = 0; m <= mmax; ++m) {
        // chunks run from l0 = max(m, |s|); the row must cover their last entry
        const int l0 = std::max(m, std::abs(p->spin));
        const int len = rup((l0 - m) + rup(lmax - l0 + 1));
        rofs[m + 1] = rofs[m] + len;
    }
    p->nrow = rofs[mmax + 1];
    auto dalloc = [&](void **ptr, size_t bytes) -> int {
        hipError_t e = hipMalloc(ptr, bytes ? bytes : 16);
        if (e != hipSuccess) return set_error(HBX_ENOMEM, "hipMalloc failed for recursion tables");
        p->allocations.push_back(*ptr);
        p->device_bytes += bytes;
        return 0;
    };
    int rc;
    if ((rc = dalloc((void **)&p->d_rofs, rofs.size() * sizeof(int64_t)))) return rc;
    // TAILPAD zero entries after the last row: chunked kernels may read a
    // few coefficients past l = lmax of the last m (they multiply zeros)
    if ((rc = dalloc((void **)&p->d_alpha, (p->nrow + TAILPAD) * sizeof(double)))) return rc;
    if ((rc = dalloc((void **)&p->d_sl, (p->nrow + TAILPAD) * sizeof(double)))) return rc;
    HBX_HIP(hipMemset(p->d_alpha + p->nrow, 0, TAILPAD * sizeof(double)));
    HBX_HIP(hipMemset(p->d_sl + p->nrow, 0, TAILPAD * sizeof(double)));
    const int64_t nseed = (int64_t)(mmax + 1) * np;
    if ((rc = dalloc((void **)&p->d_lseed, nseed * sizeof(int32_t)))) return rc;
    HBX_HIP(hipMemcpy(p->d_rofs, rofs.data(), rofs.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    if (p->spin != 0) return plan_build_spin(p, pair_cos, pair_sin, tol_log2);
    if ((rc = dalloc((void **)&p->d_seed, nseed * sizeof(double2)))) return rc;
    hipLaunchKernelGGL(recursion_tables_kernel, dim3((mmax + 1 + 127) / 128), dim3(128), 0, 0, p->d_rofs,
                       p->d_alpha, p->d_sl, lmax, mmax);
    HBX_HIP(hipGetLastError());
    // temporaries
    double *d_sin = nullptr, *d_mant = nullptr;
    int *d_ex = nullptr;
    HBX_HIP(hipMalloc(&d_sin, np * sizeof(double)));
    hipError_t e1 = hipMalloc(&d_mant, nseed * sizeof(double));
    hipError_t e2 = hipMalloc(&d_ex, nseed * sizeof(int));
    if (e1 != hipSuccess || e2 != hipSuccess) {
        (void)(void)hipFree(d_sin); if (d_mant) (void)hipFree(d_mant); if (d_ex) (void)hipFree(d_ex);
        return set_error(HBX_ENOMEM, "hipMalloc failed for seed temporaries");
    }
    hipError_t e = hipMemcpy(d_sin, pair_sin.data(), np * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(lambda_mm_kernel, dim3((np + 127) / 128), dim3(128), 0, 0, d_sin, d_mant, d_ex, np, mmax);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(seed_kernel, dim3((np + 127) / 128, mmax + 1), dim3(128), 0, 0, p->d_px, d_mant, d_ex,
                           p->d_rofs, p->d_alpha, p->d_sl, p->d_lseed, p->d_seed, np, lmax, (int)tol_log2);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(d_sin);

    (void)hipFree(d_mant);
    (void)hipFree(d_ex);
    if (e != hipSuccess) return hip_error(e, "recursion seed kernels");
    return 0;
}

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
template <typename T, int R>
__global__ void __launch_bounds__(64)
leg_analysis_kernel(const c2<T> *__restrict__ phase, int64_t pstride, c2<T> *__restrict__ alm,
                    int64_t astride, c2<T> *__restrict__ partial, const int32_t *__restrict__ pair_n,
                    const int32_t *__restrict__ pair_s, const double *__restrict__ px,
                    const int64_t *__restrict__ rofs, const double *__restrict__ alpha,
                    const double *__restrict__ sl, const int32_t *__restrict__ lseed,
                    const double2 *__restrict__ seed, int npairs, int nr, int lmax, int mmax, int ngroups,
                    int64_t nlm, int adjoint, int mbeg) {
    // transpose buffer: row v (one of the 2 CHUNK sums) of the 64 lanes, rows
    // 65 elements apart so that the reads below (4 lanes per row, 16
    // consecutive elements each) fall on distinct banks
    constexpr int RED_ROW = 65;
    __shared__ T red[2 * CHUNK * RED_ROW];
    int mb, item;
    xcd_map(blockIdx.x, MB * ngroups, &mb, &item);
    const int m = mb * MB + item / ngroups;
    const int grp = item % ngroups;
    if (m > mmax || m < mbeg) return;
    const int b = blockIdx.y;
    const int lane = threadIdx.x;
    const c2<T> *ph = phase + (int64_t)b * pstride + (int64_t)(m >> 3) * nr * MB + (m & 7);

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
            x[k] = px[p];
            ls[k] = lseed[(int64_t)m * npairs + p];
            const int rn = pair_n[p], rs = pair_s[p];
            const c2<T> fn = ph[(int64_t)rn * MB];
            const c2<T> fs = (rs != rn) ? ph[(int64_t)rs * MB] : mk<T>(0, 0);
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

    const bool direct = (ngroups == 1);
    c2<T> *dst = direct ? alm + (int64_t)b * astride : partial + ((int64_t)b * ngroups + grp) * nlm;
    const int64_t i0 = alm_index(m, m, lmax);
    const int64_t ro = rofs[m];
    // rows below every slot's switch-on are zero
    for (int l = m + lane; l < (lstart > lmax ? lmax + 1 : lstart); l += 64) dst[i0 + (l - m)] = mk<T>(0, 0);

    const int v_id = lane >> 2, q0 = (lane & 3) * 16;
    for (int l0 = lstart; l0 <= lmax; l0 += CHUNK) {
        double a[CHUNK];
#pragma unroll
        for (int i = 0; i < CHUNK; ++i) a[i] = alpha[ro + (l0 - m) + i];
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
                    const double2 sd = seed[(int64_t)m * npairs + p0 + 64 * k];
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
        for (int v = 0; v < 2 * CHUNK; ++v) red[v * RED_ROW + lane] = acc[v];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        T s = (T)0;
#pragma unroll
        for (int q = 0; q < 16; ++q) s += red[v_id * RED_ROW + q0 + q];
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
                reinterpret_cast<T *>(dst + i0 + (l - m))[c] = val;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// alm[idx(l,m)] = s_lm * sum_g partial[g][idx(l,m)] (fixed summation order);
// with lfirst, group g's partial is read only for l >= lfirst[m][g] (zero below)
template <typename T>
__global__ void group_reduce_kernel(const c2<T> *__restrict__ partial, c2<T> *__restrict__ alm, int64_t astride,
                                    const int64_t *__restrict__ rofs, const double *__restrict__ sl, int lmax,
                                    int ngroups, int64_t nlm, int adjoint, int mbeg,
                                    const int32_t *__restrict__ lfirst) {
    const int m = mbeg + blockIdx.x, b = blockIdx.y;
    const int64_t i0 = alm_index(m, m, lmax);
    const int64_t ro = rofs[m];
    const c2<T> *src = partial + (int64_t)b * ngroups * nlm + i0;
    c2<T> *dst = alm + (int64_t)b * astride + i0;
    for (int j = threadIdx.x; j <= lmax - m; j += blockDim.x) {
        double re = 0.0, im = 0.0;
        for (int g = 0; g < ngroups; ++g) {
            if (lfirst && m + j < lfirst[(int64_t)m * ngroups + g]) continue;
            const c2<T> v = src[(int64_t)g * nlm + j];
            re += (double)v.x;
            im += (double)v.y;
        }
        double s = sl[ro + j];
        if (adjoint) {
            if (m > 0) s *= 2.0;
            else im = 0.0;
        }
        dst[j] = mk<T>((T)(re * s), (T)(im * s));
    }
}

// ---------------------------------------------------------------------------
// synthesis pre-pass: scaled[row m][l-m] = s_lm * a_lm (halved for m > 0 in
// the map2alm adjoint), zero-padded rows so chunks never read past a row.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void prescale_kernel(const c2<T> *__restrict__ alm, int64_t astride, c2<T> *__restrict__ scaled,
                                int64_t nrow, const int64_t *__restrict__ rofs, const double *__restrict__ sl,
                                int lmax, int adjoint) {
    const int m = blockIdx.x, b = blockIdx.y;
    const int64_t o = rofs[m];
    const int len = (int)(rofs[m + 1] - o);
    const int64_t i0 = alm_index(m, m, lmax);
    const c2<T> *a = alm + (int64_t)b * astride + i0;
    c2<T> *d = scaled + (int64_t)b * nrow + o;
    const double f = (adjoint && m > 0) ? 0.5 : 1.0;
    for (int j = threadIdx.x; j < len; j += blockDim.x) {
        c2<T> v = mk<T>(0, 0);
        if (m + j <= lmax) v = cscale<T>(a[j], (T)(sl[o + j] * f));
        d[j] = v;
    }
}

// ---------------------------------------------------------------------------
// synthesis: scaled alm -> phase.  One wave per (m, 64 ring pairs); a
// workgroup holds the 8 m of one m-block so its writes fill whole
// [ring][m%8] lines of the phase buffer.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(512)
leg_synthesis_kernel(const c2<T> *__restrict__ scaled, int64_t nrow, c2<T> *__restrict__ phase, int64_t pstride,
                     const int32_t *__restrict__ pair_n, const int32_t *__restrict__ pair_s,
                     const double *__restrict__ px, const int64_t *__restrict__ rofs,
                     const double *__restrict__ alpha, const int32_t *__restrict__ lseed,
                     const double2 *__restrict__ seed, int npairs, int nr, int lmax, int mmax, int nmb, int npb) {
    int mb, pb;
    xcd_map(blockIdx.x, npb, &mb, &pb);
    if (mb >= nmb) return;
    const int b = blockIdx.y;
    const int w = uniform(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int m = mb * MB + w;
    if (m > mmax) return;
    const int p = pb * 64 + lane;
    const bool valid = p < npairs;
    const double x = valid ? px[p] : 0.0;
    const int ls = valid ? lseed[(int64_t)m * npairs + p] : NEVER;
    const int lstart = wave_min(ls);
    const int linj = wave_max(ls <= lmax ? ls : -1);
    c2<T> ge = mk<T>(0, 0), go = mk<T>(0, 0);
    if (lstart <= lmax) {
        const int64_t ro = rofs[m];
        const c2<T> *ar = scaled + (int64_t)b * nrow + ro - m;
        const double *al = alpha + ro - m;
        double u2 = 0.0, u1 = 0.0;
        for (int l0 = lstart; l0 <= lmax; l0 += CHUNK) {
            if (l0 <= linj && ls == l0) {
                const double2 sd = seed[(int64_t)m * npairs + p];
                u2 = sd.x;
                u1 = sd.y;
            }
#pragma unroll
            for (int i = 0; i < CHUNK; ++i) {
                const double u = fma(al[l0 + i] * x, u1, -u2);
                u2 = u1;
                u1 = u;
                const T uu = (T)u;
                const c2<T> av = ar[l0 + i];
                if (i & 1) {
                    go.x = fma(uu, av.x, go.x);
                    go.y = fma(uu, av.y, go.y);
                } else {
                    ge.x = fma(uu, av.x, ge.x);
                    ge.y = fma(uu, av.y, ge.y);
                }
            }
        }
    }
    if (!valid) return;
    // synthesis layout: ring pair p, north (ring pair_n[p]) and south (pair_s[p]) halves
    c2<T> *out = phase + (int64_t)b * pstride + syn_ofs(m, 0, p, npairs);
    out[0] = cadd<T>(ge, go);
    if (pair_s[p] != pair_n[p]) out[(int64_t)pairs8(npairs) * MB] = csub<T>(ge, go);
}

// ---------------------------------------------------------------------------
// alm2cl (alm_ops.py:19-32) on m-major alm
// ---------------------------------------------------------------------------
template <typename T>
__global__ void alm2cl_kernel(const c2<T> *__restrict__ alm, int64_t astride, double *cl, int64_t cstride,
                              int lmax, int mmax) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (l > lmax) return;
    const c2<T> *a = alm + (int64_t)b * astride;
    double acc = 0.0;
    const int top = min(l, mmax);
    for (int m = 0; m <= top; ++m) {
        const c2<T> v = a[alm_index(l, m, lmax)];
        const double p = (double)v.x * v.x + (double)v.y * v.y;
        acc += (m == 0) ? p : 2.0 * p;
    }
    cl[(int64_t)b * cstride + l] = acc / (2.0 * l + 1.0);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// Ring-pair slots per lane: FP64 holds 8 (E/O in registers cost 8 VGPRs
// each), FP32 holds 16; fewer when the geometry has fewer pairs.
// ring-pair slots (64 pairs each) per lane of the VALU analysis: the plan's
// tile if set (clamped to 8 for FP64, 16 for FP32), else enough to cover
// every pair with one wave up to 8 (FP64) / 4 (FP32).  Single-map map2alm,
// nside 2048 / 1024: FP64 R = 8 7.5 / 1.28 ms (R = 4: 9.5 / 1.59); FP32
// R = 4 6.6 / 1.39 ms (R = 8: 6.8 / 1.49, R = 16: 10.5 / 1.98)
template <typename T> static int analysis_slots(const Plan *p) {
    const int rmax = sizeof(T) == 8 ? 8 : 16;
    const int rdef = sizeof(T) == 8 ? 8 : 4;
    if (p->tile_valu_slots > 0) return std::min(p->tile_valu_slots, rmax);
    int r = 1;
    while (r < rdef && (int64_t)r * 64 < p->npairs) r *= 2;
    return r;
}

int analysis_slots_of(const Plan *p, bool f64) { return f64 ? analysis_slots<double>(p) : analysis_slots<float>(p); }

int analysis_groups(const Plan *p, bool f64) {
    const int r = f64 ? analysis_slots<double>(p) : analysis_slots<float>(p);
    return (int)((p->npairs + 64 * r - 1) / (64 * r));
}

// FP64 band of FP32 plans on the VALU (single maps): up to 4 slots of 64
// ring pairs per lane.  The band has only 8 rows of m, so fewer slots than
// the full FP64 analysis give it more waves; single-map FP32 map2alm with
// 1 / 2 / 4 / 8 slots: 7.27 / 6.29 / 5.86 / 6.02 ms at nside 2048, 1.41 /
// 1.25 / 1.24 / 1.33 ms at nside 1024, 0.47 / 0.45 / 0.46 / 0.50 ms at 512
static int band_slots(const Plan *p) { return std::min(4, analysis_slots<double>(p)); }
int band_groups(const Plan *p) { return (p->npairs + 64 * band_slots(p) - 1) / (64 * band_slots(p)); }

template <typename T>
static int leg_analysis_t(const Plan *p, int nmaps, const c2<T> *phase, int64_t pstride, c2<T> *alm,
                          int64_t astride, c2<T> *partial, int64_t nlm_stride, int mbeg, int mend, int adjoint,
                          hipStream_t s, int timer_id, int R) {
    if (mend < mbeg) return 0;
    const int G = (int)((p->npairs + 64 * R - 1) / (64 * R));
    const int nmb8 = (mend / MB + 1 + 7) / 8 * 8;
    dim3 grid(nmb8 * MB * G, nmaps);
#define HBX_LA(RV)                                                                                          \
    hipLaunchKernelGGL((leg_analysis_kernel<T, RV>), grid, dim3(64), 0, s, phase, pstride, alm, astride,    \
                       partial, p->d_pair_n, p->d_pair_s, p->d_px, p->d_rofs, p->d_alpha, p->d_sl,          \
                       p->d_lseed, p->d_seed, p->npairs, p->nrings, p->lmax, mend, G, nlm_stride, adjoint, mbeg)
    timer_mark(timer_id, true, s);
    switch (R) {
    case 1: HBX_LA(1); break;
    case 2: HBX_LA(2); break;
    case 4: HBX_LA(4); break;
    case 8: HBX_LA(8); break;
    case 16:
        if constexpr (sizeof(T) == 4) { HBX_LA(16); break; }
        return set_error(HBX_EUNSUPPORTED, "analysis slot count");
    default: return set_error(HBX_EUNSUPPORTED, "analysis slot count");
    }
#undef HBX_LA
    timer_mark(timer_id, false, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "leg_analysis_kernel launch");
    if (G > 1)
        return launch_group_reduce(p, nmaps, partial, alm, astride, G, nlm_stride, mbeg, mend, sizeof(T) == 8,
                                   adjoint, s);
    return 0;
}

int launch_group_reduce(const Plan *p, int nmaps, void *partial, void *alm, int64_t astride, int ngroups,
                        int64_t nlm_stride, int mbeg, int mend, bool f64, int adjoint, hipStream_t s,
                        const int32_t *lfirst) {
    dim3 grid(mend - mbeg + 1, nmaps);
    if (f64)
        hipLaunchKernelGGL(group_reduce_kernel<double>, grid, dim3(256), 0, s, (const double2 *)partial,
                           (double2 *)alm, astride, p->d_rofs, p->d_sl, p->lmax, ngroups, nlm_stride, adjoint, mbeg,
                           lfirst);
    else
        hipLaunchKernelGGL(group_reduce_kernel<float>, grid, dim3(256), 0, s, (const float2 *)partial,
                           (float2 *)alm, astride, p->d_rofs, p->d_sl, p->lmax, ngroups, nlm_stride, adjoint, mbeg,
                           lfirst);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "group_reduce_kernel launch");
    return 0;
}

// FP64 band of an FP32 plan: alm[b][i] = (float) band[b][i] for i < n
__global__ void band_to_f32_kernel(const double2 *__restrict__ band, int64_t n, float2 *__restrict__ alm,
                                   int64_t astride) {
    const int b = blockIdx.y;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 v = band[(int64_t)b * n + i];
        alm[(int64_t)b * astride + i] = make_float2((float)v.x, (float)v.y);
    }
}

int launch_leg_analysis(const Plan *p, int nmaps, const void *phase, void *alm, int64_t astride, void *partial,
                        int adjoint, hipStream_t s) {
    const int64_t pstride = phase_elems_per_map(p);
    if (p->precision == HBX_F64)
        return leg_analysis_t<double>(p, nmaps, (const double2 *)phase, pstride, (double2 *)alm, astride,
                                      (double2 *)partial, p->nlm, 0, p->mmax, adjoint, s,
                                      HBX_TIMER_LEGENDRE_ANALYSIS, analysis_slots<double>(p));
    return leg_analysis_t<float>(p, nmaps, (const float2 *)phase, pstride, (float2 *)alm, astride, (float2 *)partial,
                                 p->nlm, std::min(M_LO, p->mmax + 1), p->mmax, adjoint, s,
                                 HBX_TIMER_LEGENDRE_ANALYSIS, analysis_slots<float>(p));
}

// m < M_LO of an FP32 plan in FP64: band phase -> band alm (FP64) -> alm
int launch_leg_analysis_band(const Plan *p, int nmaps, const void *band_phase, void *alm, int64_t astride,
                             void *band_alm, void *band_partial, int adjoint, hipStream_t s) {
    const int mend = std::min(M_LO - 1, p->mmax);
    const int64_t nb = band_nlm(p);
    int rc;
    if (nmaps >= mfma_min_maps(p))
        rc = launch_band_analysis_mfma(p, nmaps, band_phase, band_alm, band_partial, adjoint, s);
    else
        rc = leg_analysis_t<double>(p, nmaps, (const double2 *)band_phase, band_phase_elems(p), (double2 *)band_alm,
                                    nb, (double2 *)band_partial, nb, 0, mend, adjoint, s, 0, band_slots(p));
    if (rc) return rc;
    return launch_band_to_f32(p, nmaps, band_alm, alm, astride, s);
}

int launch_band_to_f32(const Plan *p, int nrows, const void *band_alm, void *alm, int64_t astride, hipStream_t s) {
    const int64_t nb = band_nlm(p);
    hipLaunchKernelGGL(band_to_f32_kernel, dim3((unsigned)std::min<int64_t>((nb + 255) / 256, 1024), nrows),
                       dim3(256), 0, s, (const double2 *)band_alm, nb, (float2 *)alm, astride);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "band_to_f32_kernel launch");
    return 0;
}

int launch_prescale(const Plan *p, int nmaps, const void *alm, int64_t astride, void *scaled, int adjoint,
                    hipStream_t s) {
    dim3 grid(p->mmax + 1, nmaps);
    if (p->precision == HBX_F64)
        hipLaunchKernelGGL(prescale_kernel<double>, grid, dim3(256), 0, s, (const double2 *)alm, astride,
                           (double2 *)scaled, p->nrow, p->d_rofs, p->d_sl, p->lmax, adjoint);
    else
        hipLaunchKernelGGL(prescale_kernel<float>, grid, dim3(256), 0, s, (const float2 *)alm, astride,
                           (float2 *)scaled, p->nrow, p->d_rofs, p->d_sl, p->lmax, adjoint);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "prescale_kernel launch");
    return 0;
}

int launch_leg_synthesis(const Plan *p, int nmaps, const void *scaled, void *phase, hipStream_t s) {
    const int npb = (p->npairs + 63) / 64;
    const int nmb8 = (p->nmb + 7) / 8 * 8;
    dim3 grid(nmb8 * npb, nmaps);
    const int64_t pstride = phase_elems_per_map(p);
    timer_mark(HBX_TIMER_LEGENDRE_SYNTHESIS, true, s);
    if (p->precision == HBX_F64)
        hipLaunchKernelGGL(leg_synthesis_kernel<double>, grid, dim3(512), 0, s, (const double2 *)scaled, p->nrow,
                           (double2 *)phase, pstride, p->d_pair_n, p->d_pair_s, p->d_px, p->d_rofs, p->d_alpha,
                           p->d_lseed, p->d_seed, p->npairs, p->nrings, p->lmax, p->mmax, p->nmb, npb);
    else
        hipLaunchKernelGGL(leg_synthesis_kernel<float>, grid, dim3(512), 0, s, (const float2 *)scaled, p->nrow,
                           (float2 *)phase, pstride, p->d_pair_n, p->d_pair_s, p->d_px, p->d_rofs, p->d_alpha,
                           p->d_lseed, p->d_seed, p->npairs, p->nrings, p->lmax, p->mmax, p->nmb, npb);
    timer_mark(HBX_TIMER_LEGENDRE_SYNTHESIS, false, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "leg_synthesis_kernel launch");
    return 0;
}

int launch_alm2cl(int lmax, int mmax, int nmaps, const void *alm, int dt, int64_t astride, double *cl,
                  int64_t cstride, hipStream_t s) {
    dim3 grid((lmax + 1 + 255) / 256, nmaps);
    if (dt == HBX_DT_F64)
        hipLaunchKernelGGL(alm2cl_kernel<double>, grid, dim3(256), 0, s, (const double2 *)alm, astride, cl, cstride,
                           lmax, mmax);
    else
        hipLaunchKernelGGL(alm2cl_kernel<float>, grid, dim3(256), 0, s, (const float2 *)alm, astride, cl, cstride,
                           lmax, mmax);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "alm2cl_kernel launch");
    return 0;
}

}  // namespace hbx
