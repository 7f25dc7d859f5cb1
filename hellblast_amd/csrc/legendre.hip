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
#include "legendre_valu.h"

namespace hbx {


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

template <typename T, int R>
__global__ void __launch_bounds__(64)
leg_analysis_kernel(const c2<T> *__restrict__ phase, int64_t pstride, c2<T> *__restrict__ alm,
                    int64_t astride, c2<T> *__restrict__ partial, const int32_t *__restrict__ pair_n,
                    const int32_t *__restrict__ pair_s, const double *__restrict__ px,
                    const int64_t *__restrict__ rofs, const double *__restrict__ alpha,
                    const double *__restrict__ sl, const int32_t *__restrict__ lseed,
                    const double2 *__restrict__ seed, int npairs, int nr, int lmax, int mmax, int ngroups,
                    int64_t nlm, int adjoint, int mbeg) {
    __shared__ T red[red_elems<T>()];
    int mb, item;
    xcd_map(blockIdx.x, MB * ngroups, &mb, &item);
    const int m = mb * MB + item / ngroups;
    const int grp = item % ngroups;
    if (m > mmax || m < mbeg) return;
    const int b = blockIdx.y;
    const bool direct = (ngroups == 1);
    c2<T> *dst = direct ? alm + (int64_t)b * astride : partial + ((int64_t)b * ngroups + grp) * nlm;
    ana_wave<T, R>(phase + (int64_t)b * pstride + (int64_t)(m >> 3) * nr * MB + (m & 7), dst, direct, pair_n, pair_s,
                   px, rofs, alpha, sl, lseed, seed, npairs, lmax, m, grp, adjoint, red, threadIdx.x);
}

// One workgroup per (m, map); the maps of one m run on one XCD, so the
// row's s_lm (read by every map) comes into that XCD's L2 once.
template <typename T>
__global__ void group_reduce_kernel(const c2<T> *__restrict__ partial, c2<T> *__restrict__ alm, int64_t astride,
                                    const int64_t *__restrict__ rofs, const double *__restrict__ sl, int lmax,
                                    int ngroups, int64_t nlm, int adjoint, int mbeg, int nrows, int nmaps,
                                    const int32_t *__restrict__ lfirst) {
    int mi, b;
    xcd_map(blockIdx.x, nmaps, &mi, &b);
    if (mi >= nrows) return;
    const int m = mbeg + mi;
    const int64_t i0 = alm_index(m, m, lmax);
    reduce_row<T>(partial + (int64_t)b * ngroups * nlm + i0, nlm, ngroups, alm + (int64_t)b * astride + i0,
                  sl + rofs[m], lmax, m, adjoint, lfirst ? lfirst + (int64_t)m * ngroups : nullptr);
}

template <typename T>
__global__ void prescale_kernel(const c2<T> *__restrict__ alm, int64_t astride, c2<T> *__restrict__ scaled,
                                int64_t nrow, const int64_t *__restrict__ rofs, const double *__restrict__ sl,
                                int lmax, int adjoint) {
    const int m = blockIdx.x, b = blockIdx.y;
    const int64_t o = rofs[m];
    prescale_row<T>(alm + (int64_t)b * astride + alm_index(m, m, lmax), scaled + (int64_t)b * nrow + o, sl + o,
                    (int)(rofs[m + 1] - o), lmax, m, adjoint);
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
    const int64_t ro = rofs[m];
    const c2<T> *ar[1] = {scaled + (int64_t)b * nrow + ro - m};
    c2<T> *out[1] = {phase + (int64_t)b * pstride + syn_ofs(m, 0, p < npairs ? p : 0, npairs)};
    syn_wave<T, 1>(ar, out, pair_n, pair_s, px, alpha + ro - m, lseed, seed, npairs, lmax, m, p);
}

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
    const int nrows = mend - mbeg + 1;
    dim3 grid((unsigned)((nrows + 7) / 8) * 8 * (unsigned)nmaps);
    if (f64)
        hipLaunchKernelGGL(group_reduce_kernel<double>, grid, dim3(256), 0, s, (const double2 *)partial,
                           (double2 *)alm, astride, p->d_rofs, p->d_sl, p->lmax, ngroups, nlm_stride, adjoint, mbeg,
                           nrows, nmaps, lfirst);
    else
        hipLaunchKernelGGL(group_reduce_kernel<float>, grid, dim3(256), 0, s, (const float2 *)partial,
                           (float2 *)alm, astride, p->d_rofs, p->d_sl, p->lmax, ngroups, nlm_stride, adjoint, mbeg,
                           nrows, nmaps, lfirst);
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
