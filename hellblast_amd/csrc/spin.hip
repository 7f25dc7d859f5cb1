// Spin-weighted Legendre stage (spin s >= 1): Q/U <-> E/B.
//
// Reference: spin.py:32-115 (map2alm_spin_gpu, eb_from_qu_gpu, qu_from_eb_gpu)
// and spin1.py:5-69, whose Wigner kernel (wigner_d_ms.cu:20-50) costs O(l) per
// element and uses wrong exponents.  Here sY_lm = a_lm(theta) e^{i m phi}
// (Goldberg et al. 1967), a = s_lambda, b = (-s)_lambda, sigma = (-1)^s, and
// with P = F_Q + i F_U, M = F_Q - i F_U (ring coefficients, weights included)
//   S_a = sum_r a P,  S_b = sigma sum_r b M,
//   E = -(S_a + S_b)/2,  B = i (S_a - S_b)/2            (spin.py:73-78)
// and for synthesis, with u = E + iB, v = E - iB,
//   P(m) = -sum_l a u,  M(m) = -sigma sum_l b v,
//   G_Q = (P + M)/2,    G_U = -i (P - M)/2.
// a and b come from their own Wigner-d l-recursions,
//   a_{l+1} = A_l (x + B_l) a_l - C_l a_{l-1},  B_l = m s/(l(l+1)),  b: B_l -> -B_l,
// rescaled (lambda = s_l mu, s_{l+1} = C_l s_{l-1}) to
//   mu_a(l) = (alpha_l x + beta_l) mu_a(l-1) - mu_a(l-2),  mu_b: beta_l -> -beta_l.
// Two independent recursions keep each function's relative accuracy near the
// poles, where a and b differ by tan^{2s}(theta/2) (a recursion for the
// combinations (a +- sigma b)/2 loses it).  Mirror rings: a(pi - theta) =
// (-1)^{l+m} b(theta), so one recursion serves a ring pair.
#include <algorithm>
#include <cmath>

#include "hbx_device.h"
#include "hbx_internal.h"

namespace hbx {

constexpr int SNEVER = 0x7fffffff;

// ---------------------------------------------------------------------------
// plan-time tables
// ---------------------------------------------------------------------------
__global__ void spin_tables_kernel(const int64_t *__restrict__ rofs, double *alpha, double *beta, double *sl,
                                   int lmax, int mmax, int s) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m > mmax) return;
    const int64_t o = rofs[m];
    const int len = (int)(rofs[m + 1] - o);
    const int l0 = max(m, abs(s));
    const double dm = m, ds = s;
    double s2 = 1.0, s1 = 1.0;   // s_{l-2}, s_{l-1}
    for (int j = 0; j < len; ++j) {
        const int l = m + j;                  // target degree of this entry
        double a = 0.0, b = 0.0, sv = 0.0;
        if (l >= l0 && l <= lmax) {
            if (l == l0) {
                sv = 1.0;
            } else {
                const int k = l - 1;          // recursion from k to k+1 = l
                const double dk = k;
                const double den = sqrt(((dk + 1) * (dk + 1) - dm * dm) * ((dk + 1) * (dk + 1) - ds * ds));
                const double A = (dk + 1) * sqrt((2 * dk + 1) * (2 * dk + 3)) / den;
                const double Bk = dm * ds / (dk * (dk + 1));
                if (l == l0 + 1) {
                    sv = 1.0;
                } else {
                    const double C = (dk + 1) * sqrt((dk * dk - dm * dm) * (dk * dk - ds * ds)) *
                                     sqrt((2 * dk + 3) / (2 * dk - 1)) / (dk * den);
                    sv = C * s2;
                }
                a = A * s1 / sv;
                b = A * Bk * s1 / sv;
            }
            s2 = s1;
            s1 = sv;
        }
        alpha[o + j] = a;
        beta[o + j] = b;
        sl[o + j] = sv;
    }
}

// s_lambda and (-s)_lambda at l0 = max(m, |s|) as (mant_a, mant_b) * 2^ex, one
// thread per ring pair
__global__ void spin_l0_kernel(const double *__restrict__ pcos, const double *__restrict__ psin, double *mp,
                               double *mm, int *ex, int npairs, int mmax, int s) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    const double x = pcos[p], st = psin[p];
    // half angles from x and sin theta without cancellation
    const double sh = (x >= 0.0) ? st / sqrt(2.0 * (1.0 + x)) : sqrt(0.5 * (1.0 - x));
    const double ch = (x >= 0.0) ? sqrt(0.5 * (1.0 + x)) : st / sqrt(2.0 * (1.0 - x));
    const int a = abs(s);
    auto fact = [](int n) { double f = 1.0; for (int i = 2; i <= n; ++i) f *= i; return f; };
    auto comb = [&](int n, int k) { return fact(n) / (fact(k) * fact(n - k)); };
    auto store = [&](int m, double vp, int ep, double vm, int em) {
        // s_lambda and (-s)_lambda with a common exponent
        const int e = max(ep, em);
        const int64_t q = (int64_t)m * npairs + p;
        mp[q] = ldexp(vp, ep - e);
        mm[q] = ldexp(vm, em - e);
        ex[q] = e;
    };
    // m < |s|: closed forms at l0 = |s|
    const double base0 = 1.0 / sqrt(4.0 * 3.14159265358979323846);
    for (int m = 0; m < a && m <= mmax; ++m) {
        const double base = base0 * sqrt(fact(a + m) * fact(a - m) * (2 * a + 1) / fact(2 * a));
        const double vs = ((m & 1) ? -1.0 : 1.0) * base * comb(2 * a, a - m) * pow(sh, a + m) * pow(ch, a - m);
        const double vms = ((a & 1) ? -1.0 : 1.0) * base * comb(2 * a, a + m) * pow(sh, a - m) * pow(ch, a + m);
        int e1 = 0, e2 = 0;
        const double f1 = frexp(vs, &e1), f2 = frexp(vms, &e2);
        store(m, f1, e1, f2, e2);
    }
    if (mmax < a) return;
    // m >= |s|: s_lambda_mm chains for +s and -s, renormalised with frexp
    int ep = 0, em = 0;
    double vp = frexp(((a & 1) ? -1.0 : 1.0) * sqrt((2 * a + 1) * base0 * base0), &ep);
    double vm = vp;
    em = ep;
    for (int i = 0; i < 2 * a; ++i) {
        int d1, d2;
        vp = frexp(vp * sh, &d1);
        ep += d1;
        vm = frexp(vm * ch, &d2);
        em += d2;
    }
    store(a, vp, ep, vm, em);
    for (int m = a; m < mmax; ++m) {
        const double r1 = -sqrt((2.0 * m + 3) * (2.0 * m + 2) / ((double)(m + 1 + s) * (double)(m + 1 - s))) * 0.5 * st;
        const double r2 = -sqrt((2.0 * m + 3) * (2.0 * m + 2) / ((double)(m + 1 - s) * (double)(m + 1 + s))) * 0.5 * st;
        int d1, d2;
        vp = frexp(vp * r1, &d1);
        ep += d1;
        vm = frexp(vm * r2, &d2);
        em += d2;
        store(m + 1, vp, ep, vm, em);
    }
}

// switch-on seeds of the coupled recursion, one thread per (m, ring pair)
__global__ void spin_seed_kernel(const double *__restrict__ px, const double *__restrict__ mp0,
                                 const double *__restrict__ mm0, const int *__restrict__ ex0,
                                 const int64_t *__restrict__ rofs, const double *__restrict__ alpha,
                                 const double *__restrict__ beta, const double *__restrict__ sl, int32_t *lseed,
                                 double4 *seed, int npairs, int lmax, int s, int tol_log2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = blockIdx.y;
    if (p >= npairs) return;
    const int64_t q = (int64_t)m * npairs + p;
    const int l0 = max(m, abs(s));
    const double x = px[p];
    const double a0 = mp0[q], b0 = mm0[q];
    const int e0 = ex0[q];
    const double thr = ldexp(1.0, tol_log2);
    const double big = ldexp(1.0, 200), small = ldexp(1.0, -200);
    const double *al = alpha + rofs[m] - m, *be = beta + rofs[m] - m, *sv = sl + rofs[m] - m;
    int lon = SNEVER;
    if (l0 <= lmax && (a0 != 0.0 || b0 != 0.0)) {
        double p2 = 0.0, p1 = a0, m2 = 0.0, m1 = b0;
        int e = e0;
        if (fmax(fabs(ldexp(p1, e)), fabs(ldexp(m1, e))) >= thr) {
            lon = l0;
        } else {
            for (int l = l0 + 1; l <= lmax; ++l) {
                const double t = al[l] * x;
                const double np = fma(t + be[l], p1, -p2);
                const double nm = fma(t - be[l], m1, -m2);
                p2 = p1; p1 = np; m2 = m1; m1 = nm;
                if (fmax(fabs(p1), fabs(m1)) > big) { p1 *= small; p2 *= small; m1 *= small; m2 *= small; e += 200; }
                if (e > tol_log2 - 1100 && fmax(fabs(ldexp(sv[l] * p1, e)), fabs(ldexp(sv[l] * m1, e))) >= thr) {
                    lon = l;
                    break;
                }
            }
        }
    }
    if (lon == SNEVER) {
        lseed[q] = SNEVER;
        seed[q] = make_double4(0, 0, 0, 0);
        return;
    }
    const int ls = l0 + ((lon - l0) / CHUNK) * CHUNK;
    lseed[q] = ls;
    if (ls == l0) {
        seed[q] = make_double4(-ldexp(a0, e0), 0.0, -ldexp(b0, e0), 0.0);
        return;
    }
    double p2 = 0.0, p1 = a0, m2 = 0.0, m1 = b0;
    int e = e0;
    for (int l = l0 + 1; l <= ls - 1; ++l) {
        const double t = al[l] * x;
        const double np = fma(t + be[l], p1, -p2);
        const double nm = fma(t - be[l], m1, -m2);
        p2 = p1; p1 = np; m2 = m1; m1 = nm;
        if (fmax(fabs(p1), fabs(m1)) > big) { p1 *= small; p2 *= small; m1 *= small; m2 *= small; e += 200; }
    }
    seed[q] = make_double4(ldexp(p2, e), ldexp(p1, e), ldexp(m2, e), ldexp(m1, e));
}

int plan_build_spin(Plan *p, const std::vector<double> &pair_cos, const std::vector<double> &pair_sin,
                    double tol_log2) {
    const int mmax = p->mmax, lmax = p->lmax, np = p->npairs, s = p->spin;
    auto dalloc = [&](void **ptr, size_t bytes) -> int {
        hipError_t e = hipMalloc(ptr, bytes ? bytes : 16);
        if (e != hipSuccess) return set_error(HBX_ENOMEM, "hipMalloc failed for spin tables");
        p->allocations.push_back(*ptr);
        p->device_bytes += bytes;
        return 0;
    };
    int rc;
    if ((rc = dalloc((void **)&p->d_beta, (p->nrow + TAILPAD) * sizeof(double)))) return rc;
    HBX_HIP(hipMemset(p->d_beta + p->nrow, 0, TAILPAD * sizeof(double)));
    const int64_t nseed = (int64_t)(mmax + 1) * np;
    if ((rc = dalloc((void **)&p->d_seed4, nseed * sizeof(double4)))) return rc;
    hipLaunchKernelGGL(spin_tables_kernel, dim3((mmax + 1 + 127) / 128), dim3(128), 0, 0, p->d_rofs, p->d_alpha,
                       p->d_beta, p->d_sl, lmax, mmax, s);
    HBX_HIP(hipGetLastError());
    double *d_cos = nullptr, *d_sin = nullptr, *d_mp = nullptr, *d_mm = nullptr;
    int *d_ex = nullptr;
    hipError_t e = hipMalloc(&d_cos, np * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&d_sin, np * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&d_mp, nseed * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&d_mm, nseed * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&d_ex, nseed * sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(d_cos, pair_cos.data(), np * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_sin, pair_sin.data(), np * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(spin_l0_kernel, dim3((np + 127) / 128), dim3(128), 0, 0, d_cos, d_sin, d_mp, d_mm, d_ex, np,
                           mmax, s);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(spin_seed_kernel, dim3((np + 127) / 128, mmax + 1), dim3(128), 0, 0, p->d_px, d_mp, d_mm,
                           d_ex, p->d_rofs, p->d_alpha, p->d_beta, p->d_sl, p->d_lseed, p->d_seed4, np, lmax, s,
                           (int)tol_log2);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    for (void *ptr : {(void *)d_cos, (void *)d_sin, (void *)d_mp, (void *)d_mm, (void *)d_ex})
        if (ptr) (void)hipFree(ptr);
    if (e != hipSuccess) return hip_error(e, "spin seed kernels");
    return 0;
}

// ---------------------------------------------------------------------------
// analysis: phase of (Q, U) -> (E, B).  One wave per (m, ring-pair group,
// pair of maps); R slots of 64 ring pairs per lane; 32 partial sums per
// 8-l chunk reduced over the wave through LDS.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int swmin(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int swmax(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return __builtin_amdgcn_readfirstlane(v);
}

template <typename T, int R>
__global__ void __launch_bounds__(64)
spin_analysis_kernel(const c2<T> *__restrict__ phase, int64_t pstride, c2<T> *__restrict__ alm, int64_t astride,
                     c2<T> *__restrict__ partial, const double *__restrict__ px, const int64_t *__restrict__ rofs,
                     const double *__restrict__ alpha, const double *__restrict__ beta,
                     const double *__restrict__ sl, const int32_t *__restrict__ lseed,
                     const double4 *__restrict__ seed, int npairs, int nr, int lmax, int mmax, int ngroups,
                     int64_t nlm, int s, int adjoint) {
    __shared__ T red[4 * CHUNK * 64];
    int mb, item;
    xcd_map(blockIdx.x, MB * ngroups, &mb, &item);
    const int m = mb * MB + item / ngroups;
    const int grp = item % ngroups;
    if (m > mmax) return;
    const int b = blockIdx.y;                 // Q/U pair index
    const int lane = threadIdx.x;
    const int l0 = max(m, abs(s));
    const int par0 = (l0 + m) & 1;            // parity of l + m at the first l of every chunk
    const c2<T> *phq = phase + (int64_t)(2 * b) * pstride + (int64_t)(m >> 3) * nr * MB + (m & 7);
    const c2<T> *phu = phq + pstride;

    // per slot: P_n = Q_n + i U_n, M_n = Q_n - i U_n; south P_s, M_s carry the
    // sign (-1)^{l+m} of the even-i steps of a chunk
    double x[R], p1[R], p2[R], q1[R], q2[R];
    c2<T> Pn[R], Mn[R], Ps[R], Ms[R];
    int ls[R], sst[R], sinj[R];
    const int pb = grp * 64 * R + lane;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int p = pb + 64 * k;
        p1[k] = p2[k] = q1[k] = q2[k] = 0.0;
        if (p < npairs) {
            x[k] = px[p];
            ls[k] = lseed[(int64_t)m * npairs + p];
            const int rs = nr - 1 - p;
            const c2<T> qn = phq[(int64_t)p * MB], un = phu[(int64_t)p * MB];
            const c2<T> qs = (rs != p) ? phq[(int64_t)rs * MB] : mk<T>(0, 0);
            const c2<T> us = (rs != p) ? phu[(int64_t)rs * MB] : mk<T>(0, 0);
            const T sg = par0 ? (T)-1 : (T)1;
            Pn[k] = mk<T>(qn.x - un.y, qn.y + un.x);
            Mn[k] = mk<T>(qn.x + un.y, qn.y - un.x);
            Ps[k] = mk<T>(sg * (qs.x - us.y), sg * (qs.y + us.x));
            Ms[k] = mk<T>(sg * (qs.x + us.y), sg * (qs.y - us.x));
        } else {
            x[k] = 0.0;
            ls[k] = SNEVER;
            Pn[k] = Mn[k] = Ps[k] = Ms[k] = mk<T>(0, 0);
        }
        sst[k] = swmin(ls[k]);
        sinj[k] = swmax(ls[k] <= lmax ? ls[k] : -1);
    }
    int lstart = SNEVER;
#pragma unroll
    for (int k = 0; k < R; ++k) lstart = min(lstart, sst[k]);

    const bool direct = (ngroups == 1);
    c2<T> *dE = direct ? alm + (int64_t)(2 * b) * astride : partial + ((int64_t)(2 * b) * ngroups + grp) * nlm;
    c2<T> *dB = direct ? alm + (int64_t)(2 * b + 1) * astride : partial + ((int64_t)(2 * b + 1) * ngroups + grp) * nlm;
    const int64_t i0 = alm_index(m, m, lmax) - m;    // + l
    const int64_t ro = rofs[m] - m;                   // table index of l
    for (int l = m + lane; l < (lstart > lmax ? lmax + 1 : lstart); l += 64) {
        dE[i0 + l] = mk<T>(0, 0);
        dB[i0 + l] = mk<T>(0, 0);
    }
    const double sigma = (s & 1) ? -1.0 : 1.0;
    const int vid = lane >> 1, half = lane & 1;
    for (int c0 = lstart; c0 <= lmax; c0 += CHUNK) {
        double al[CHUNK], be[CHUNK];
#pragma unroll
        for (int i = 0; i < CHUNK; ++i) { al[i] = alpha[ro + c0 + i]; be[i] = beta[ro + c0 + i]; }
        T acc[4 * CHUNK];     // per l: S_a.re, S_a.im, S_b.re, S_b.im (S_b without sigma)
#pragma unroll
        for (int i = 0; i < 4 * CHUNK; ++i) acc[i] = (T)0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (c0 < sst[k]) continue;
            if (c0 <= sinj[k] && ls[k] == c0) {
                const double4 sd = seed[(int64_t)m * npairs + pb + 64 * k];
                p2[k] = sd.x; p1[k] = sd.y; q2[k] = sd.z; q1[k] = sd.w;
            }
#pragma unroll
            for (int i = 0; i < CHUNK; ++i) {
                const double t = al[i] * x[k];
                const double na = fma(t + be[i], p1[k], -p2[k]);
                const double nb = fma(t - be[i], q1[k], -q2[k]);
                p2[k] = p1[k]; p1[k] = na; q2[k] = q1[k]; q1[k] = nb;
                const T a = (T)na;
                const T bs = (i & 1) ? -(T)nb : (T)nb;    // (-1)^{l+m} relative to the chunk start
                const T as = (i & 1) ? -a : a;
                acc[4 * i + 0] = fma(a, Pn[k].x, fma(bs, Ps[k].x, acc[4 * i + 0]));
                acc[4 * i + 1] = fma(a, Pn[k].y, fma(bs, Ps[k].y, acc[4 * i + 1]));
                acc[4 * i + 2] = fma((T)nb, Mn[k].x, fma(as, Ms[k].x, acc[4 * i + 2]));
                acc[4 * i + 3] = fma((T)nb, Mn[k].y, fma(as, Ms[k].y, acc[4 * i + 3]));
            }
        }
#pragma unroll
        for (int v = 0; v < 4 * CHUNK; ++v) red[v * 64 + lane] = acc[v];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        T sum = (T)0;
#pragma unroll
        for (int q = 0; q < 32; ++q) sum += red[vid * 64 + half * 32 + ((q + vid) & 31)];
        sum += __shfl_xor(sum, 1, 64);
        // partner: the same l's S_b component for an S_a lane and vice versa
        const T partner = __shfl_xor(sum, 4, 64);
        if (half == 0) {
            const int i = vid >> 2, comp = vid & 3;
            const int l = c0 + i;
            if (l <= lmax) {
                // comp 0/1: E.re/E.im = -(S_a + sigma S_b)/2;
                // comp 2: B.im = (S_a.re - sigma S_b.re)/2; comp 3: B.re = -(S_a.im - sigma S_b.im)/2
                double v;
                if (comp < 2) v = -0.5 * ((double)sum + sigma * (double)partner);
                else if (comp == 2) v = 0.5 * ((double)partner - sigma * (double)sum);
                else v = -0.5 * ((double)partner - sigma * (double)sum);
                const bool imag = (comp == 1 || comp == 2);
                if (direct) {
                    v *= sl[ro + l];
                    if (adjoint) {
                        if (m > 0) v *= 2.0;
                        else if (imag) v = 0.0;
                    }
                }
                T *dst = reinterpret_cast<T *>((comp < 2 ? dE : dB) + i0 + l);
                dst[imag ? 1 : 0] = (T)v;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// ---------------------------------------------------------------------------
// analysis on the matrix cores (FP32 plans, m >= M_LO).  Per m the sums of
// spin_analysis_kernel,
//   S_a(l) = sum_p a_l(p) P_n(p) + eps b_l(p) P_s(p),
//   S_b(l) = sum_p b_l(p) M_n(p) + eps a_l(p) M_s(p),     eps = (-1)^{l+m},
// are two products over the ring pairs p.  With 16 output rows n = 4 j + c
// (polarisation pair j of a group of 4, c = S_a.re, S_a.im, S_b.re, S_b.im)
//   D[n][l] = sum_p A_a[n][p] a_l(p) + A_b[n][p] b_l(p),
//   A_a[n][p] = (P_n.re, P_n.im, eps M_s.re, eps M_s.im)[c],
//   A_b[n][p] = (eps P_s.re, eps P_s.im, M_n.re, M_n.im)[c],
// on v_mfma_f32_16x16x4f32: FP32 operands (the FP32 values the VALU kernel
// multiplies) and FP32 sums.  A workgroup is 4 waves over P = 128 ring pairs
// of one m.  Per 32-l chunk waves 0-1 run the FP64 a-recursion and waves 2-3
// the b-recursion of the 128 pairs (one pair per lane) into two lambda tiles;
// then wave (par, fn) contracts A_fn with tile fn over the l of parity par
// (chunks start at l = m mod 2, so eps = (-1)^par), and the b-waves hand
// their D tiles to the a-waves through LDS.  Tiles as in the scalar matrix
// kernel: [par][col][k][s], pair 4 s + k.
// Lane layout of the 16x16x4 f32 instruction (checked on gfx950):
//   A: lane = row + 16 k   B: lane = col + 16 k   D: lane = col + 16 (row/4), reg = row%4
// ---------------------------------------------------------------------------
typedef float sfloatx4 __attribute__((ext_vector_type(4)));

constexpr int SPL = 32;                   // l per chunk: 16 even + 16 odd columns

struct SpinMfGeom {
    static constexpr int W = 4;                        // waves per workgroup
    static constexpr int P = 128;                      // ring pairs per workgroup
    static constexpr int RUN = P / 4 + 16;             // stride between the k runs of a column
    static constexpr int COL = P + 68;                 // column stride (= 4 mod 64 banks)
    static constexpr int PAR = 16 * COL;
    static constexpr int TILE = 2 * PAR;               // floats per lambda function
    static constexpr int XCH = 2 * 4 * 64;             // b-waves -> a-waves: parity x reg x lane
    static constexpr size_t LDS = (size_t)(2 * TILE + XCH) * sizeof(float) + 2 * 2 * SPL * sizeof(double) + 16;
};

__global__ void __launch_bounds__(SpinMfGeom::W * 64)
spin_analysis_mfma_kernel(const float2 *__restrict__ phase, int64_t pstride, float2 *__restrict__ alm,
                          int64_t astride, float2 *__restrict__ partial, const double *__restrict__ px,
                          const int64_t *__restrict__ rofs, const double *__restrict__ alpha,
                          const double *__restrict__ beta, const double *__restrict__ sl,
                          const int32_t *__restrict__ lseed, const double4 *__restrict__ seed, int npairs, int nr,
                          int lmax, int mmax, int ngroups, int64_t nlm, int npol, int s, int adjoint, int mbeg,
                          int32_t *__restrict__ lfirst) {
    using G = SpinMfGeom;
    extern __shared__ __align__(16) float sms[];
    float *lam = sms;                                                  // [2][TILE]: a, b
    float *xch = sms + 2 * G::TILE;                                    // [XCH]
    double *cf = reinterpret_cast<double *>(xch + G::XCH);             // [2][alpha SPL | beta SPL]
    int *start = reinterpret_cast<int *>(cf + 4 * SPL);

    int mb, item;
    xcd_map(blockIdx.x, MB * ngroups, &mb, &item);
    // the 8 m of one m-block (sharing 64-byte phase lines) run next to each other
    const int m = mb * MB + item % MB;
    const int grp = item / MB;
    if (m > mmax || m < mbeg) return;
    const int g = blockIdx.y;                          // group of 4 polarisation pairs
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int par = w & 1, fn = w >> 1;                // contraction: parity; lambda function (0: a, 1: b)
    const int P0 = grp * G::P;

    // recursion: function fn of pair P0 + 64 (w & 1) + lane
    const int qr = 64 * (w & 1) + lane;
    const int pr = P0 + qr;
    double x = 0.0, u1 = 0.0, u2 = 0.0;
    double2 sd = make_double2(0.0, 0.0);
    int ls = SNEVER;
    if (pr < npairs) {
        x = px[pr];
        ls = lseed[(int64_t)m * npairs + pr];
        if (ls <= lmax) {
            const double4 q = seed[(int64_t)m * npairs + pr];
            sd = fn ? make_double2(q.z, q.w) : make_double2(q.x, q.y);
        }
    }
    const int wstart = swmin(ls);
    const int winj = swmax(ls <= lmax ? ls : -1);

    // A operand of this wave: A_fn[n = lane%16][pair P0 + 4 s + lane/16], s = 0..31
    float Sv[32];
    {
        const int row = lane & 15, k = lane >> 4;
        const int j = row >> 2, c = row & 3;
        const int gb = g * 4 + j;
        const bool south = (fn == 0) == (c >= 2);      // a: P_n, M_s; b: P_s, M_n
        const int cq = c & 1, cu = cq ^ 1;             // Q part re/im, U part im/re
        // P = Q + iU = (Q.re - U.im, Q.im + U.re), M = Q - iU = (Q.re + U.im, Q.im - U.re)
        const float su = ((c >= 2) != (cq == 1)) ? 1.0f : -1.0f;
        const float sg = (south && par) ? -1.0f : 1.0f;
        const float *pq = reinterpret_cast<const float *>(phase + (int64_t)(2 * min(gb, npol - 1)) * pstride +
                                                          (int64_t)mb * nr * MB + (m & 7));
        const float *pu = pq + 2 * pstride;
#pragma unroll
        for (int s4 = 0; s4 < 32; ++s4) {
            const int q = P0 + 4 * s4 + k;
            const int qq = min(q, npairs - 1);
            const int r = south ? nr - 1 - qq : qq;
            const bool ok = q < npairs && gb < npol && !(south && nr - 1 - qq == qq);
            const float vq = pq[(int64_t)r * (2 * MB) + cq], vu = pu[(int64_t)r * (2 * MB) + cu];
            Sv[s4] = ok ? sg * (vq + su * vu) : 0.0f;
        }
    }
    if (tid == 0) start[0] = SNEVER;
    __syncthreads();
    if (lane == 0) atomicMin(&start[0], wstart);
    __syncthreads();
    const int lstart = start[0];

    const bool direct = (ngroups == 1);
    const int64_t i0 = alm_index(m, m, lmax) - m;       // + l
    const int64_t ro = rofs[m] - m;                      // table index of l
    const double *sv = sl + ro;
    // out(l, b): b = alm row (E of pair j: 8 g + 2 j, B: 8 g + 2 j + 1)
    auto store = [&](int l, int b, double re, double im) {
        if (b >= 2 * npol) return;
        if (direct) {
            const double f = sv[l] * ((adjoint && m > 0) ? 2.0 : 1.0);
            alm[(int64_t)b * astride + i0 + l] = make_float2((float)(re * f), (float)(im * f));
        } else {
            partial[((int64_t)b * ngroups + grp) * nlm + i0 + l] = make_float2((float)re, (float)im);
        }
    };
    // coefficients below every ring pair's switch-on are zero: written to alm
    // by a single group; partial sums record where theirs start instead
    const int lz = lstart > lmax ? lmax + 1 : lstart;
    if (direct)
        for (int e = tid; e < (lz - m) * 8; e += G::W * 64) store(m + (e >> 3), g * 8 + (e & 7), 0.0, 0.0);
    else if (g == 0 && tid == 0)
        lfirst[(int64_t)m * ngroups + grp] = lz;
    if (lstart > lmax) return;

    // alpha and beta of a chunk staged in LDS (double-buffered), loaded a
    // whole chunk ahead; rows are zero past lmax within the row, and the
    // tables carry TAILPAD zeros after the last row
    auto coef = [&](int l) -> double { return tid < SPL ? alpha[ro + l + tid] : beta[ro + l + tid - SPL]; };
    if (tid < 2 * SPL) cf[tid] = coef(lstart);
    double cnext = (tid < 2 * SPL) ? coef(lstart + SPL) : 0.0;
    __syncthreads();
    // this lane's writes: pair qr -> [.][.][k = qr%4][s = qr/4] of tile fn
    float *lw = lam + fn * G::TILE + (qr & 3) * G::RUN + (qr >> 2);
    // this lane's B reads: col = lane%16, k = lane/16, s = 0..31 of tile fn
    const float *lr = lam + fn * G::TILE + par * G::PAR + (lane & 15) * G::COL + (lane >> 4) * G::RUN;
    const double sigma = (s & 1) ? -1.0 : 1.0;
    int buf = 0;
    for (int c0 = lstart; c0 <= lmax; c0 += SPL, buf ^= 1) {
        const double *ca = cf + buf * 2 * SPL;
        const double cnext2 = (tid < 2 * SPL) ? coef(c0 + 2 * SPL) : 0.0;   // in flight a chunk
        // 1. recursion of this lane's (pair, function) over l = c0 .. c0+31
#pragma unroll
        for (int tt = 0; tt < SPL; ++tt) {
            if ((tt & 7) == 0 && c0 <= winj && ls == c0 + tt) {
                u2 = sd.x;
                u1 = sd.y;
            }
            const double t = ca[tt] * x;
            const double bl = fn ? -ca[SPL + tt] : ca[SPL + tt];
            const double u = fma(t + bl, u1, -u2);
            u2 = u1;
            u1 = u;
            lw[(tt & 1) * G::PAR + (tt >> 1) * G::COL] = (float)u;
        }
        if (tid < 2 * SPL) cf[(buf ^ 1) * 2 * SPL + tid] = cnext;
        cnext = cnext2;
        __syncthreads();
        // 2. this wave's product over the 128 pairs; four chains of 32 pairs
        sfloatx4 D0 = sfloatx4{0.f, 0.f, 0.f, 0.f}, D1 = D0, D2 = D0, D3 = D0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float4 bq = *reinterpret_cast<const float4 *>(lr + 4 * j);
            sfloatx4 &Dj = (j < 2) ? D0 : (j < 4) ? D1 : (j < 6) ? D2 : D3;
            Dj = __builtin_amdgcn_mfma_f32_16x16x4f32(Sv[4 * j], bq.x, Dj, 0, 0, 0);
            Dj = __builtin_amdgcn_mfma_f32_16x16x4f32(Sv[4 * j + 1], bq.y, Dj, 0, 0, 0);
            Dj = __builtin_amdgcn_mfma_f32_16x16x4f32(Sv[4 * j + 2], bq.z, Dj, 0, 0, 0);
            Dj = __builtin_amdgcn_mfma_f32_16x16x4f32(Sv[4 * j + 3], bq.w, Dj, 0, 0, 0);
        }
        sfloatx4 D = (D0 + D1) + (D2 + D3);
        // 3. the b-waves hand their tiles to the a-waves
        if (fn == 1) {
#pragma unroll
            for (int r = 0; r < 4; ++r) xch[(par * 4 + r) * 64 + lane] = D[r];
        }
        __syncthreads();
        if (fn == 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) D[r] += xch[(par * 4 + r) * 64 + lane];
            // D[row][col]: lane = col + 16 (row/4), reg = row%4 -> pair j = lane/16,
            // (S_a.re, S_a.im, S_b.re, S_b.im) of l = c0 + 2 col + par
            const int l = c0 + 2 * (lane & 15) + par;
            if (l <= lmax) {
                const int b0 = g * 8 + 2 * (lane >> 4);
                const double sar = D[0], sai = D[1], sbr = D[2], sbi = D[3];
                // E = -(S_a + sigma S_b)/2, B = i (S_a - sigma S_b)/2
                store(l, b0, -0.5 * (sar + sigma * sbr), -0.5 * (sai + sigma * sbi));
                store(l, b0 + 1, -0.5 * (sai - sigma * sbi), 0.5 * (sar - sigma * sbr));
            }
        }
    }
}

// synthesis pre-pass: row 2b = s_l (E + iB), row 2b+1 = s_l (E - iB) (halved
// for m > 0 in the map2alm adjoint), zero-padded to the row length
template <typename T>
__global__ void spin_prescale_kernel(const c2<T> *__restrict__ alm, int64_t astride, c2<T> *__restrict__ scaled,
                                     int64_t nrow, const int64_t *__restrict__ rofs, const double *__restrict__ sl,
                                     int lmax, int adjoint) {
    const int m = blockIdx.x, b = blockIdx.y;
    const int64_t o = rofs[m];
    const int len = (int)(rofs[m + 1] - o);
    const int64_t i0 = alm_index(m, m, lmax);
    const c2<T> *E = alm + (int64_t)(2 * b) * astride + i0;
    const c2<T> *B = alm + (int64_t)(2 * b + 1) * astride + i0;
    c2<T> *u = scaled + (int64_t)(2 * b) * nrow + o;
    c2<T> *v = scaled + (int64_t)(2 * b + 1) * nrow + o;
    const double f = (adjoint && m > 0) ? 0.5 : 1.0;
    for (int j = threadIdx.x; j < len; j += blockDim.x) {
        c2<T> pu = mk<T>(0, 0), pv = pu;
        if (m + j <= lmax) {
            const T g = (T)(sl[o + j] * f);
            const c2<T> e = E[j], bb = B[j];
            pu = mk<T>(g * (e.x - bb.y), g * (e.y + bb.x));
            pv = mk<T>(g * (e.x + bb.y), g * (e.y - bb.x));
        }
        u[j] = pu;
        v[j] = pv;
    }
}

// ---------------------------------------------------------------------------
// synthesis: u, v rows -> phase of (Q, U).  One wave per (m, 64 ring pairs);
// the 8 m of an m-block share a workgroup.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(512)
spin_synthesis_kernel(const c2<T> *__restrict__ scaled, int64_t nrow, c2<T> *__restrict__ phase, int64_t pstride,
                      const double *__restrict__ px, const int64_t *__restrict__ rofs,
                      const double *__restrict__ alpha, const double *__restrict__ beta,
                      const int32_t *__restrict__ lseed, const double4 *__restrict__ seed, int npairs, int nr,
                      int lmax, int mmax, int nmb, int npb, int s) {
    int mb, pb;
    xcd_map(blockIdx.x, npb, &mb, &pb);
    if (mb >= nmb) return;
    const int b = blockIdx.y;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int m = mb * MB + w;
    if (m > mmax) return;
    const int p = pb * 64 + lane;
    const bool valid = p < npairs;
    const int l0 = max(m, abs(s));
    const int par0 = (l0 + m) & 1;
    const double x = valid ? px[p] : 0.0;
    const int ls = valid ? lseed[(int64_t)m * npairs + p] : SNEVER;
    const int lstart = swmin(ls);
    const int linj = swmax(ls <= lmax ? ls : -1);
    // sums by class (A: even step of a chunk, B: odd): a u, b v (north); b u, a v (south)
    c2<T> auA = mk<T>(0, 0), auB = auA, bvA = auA, bvB = auA, buA = auA, buB = auA, avA = auA, avB = auA;
    if (lstart <= lmax) {
        const int64_t ro = rofs[m] - m;
        const c2<T> *ru = scaled + (int64_t)(2 * b) * nrow + ro;
        const c2<T> *rv = scaled + (int64_t)(2 * b + 1) * nrow + ro;
        double p2 = 0.0, p1 = 0.0, q2 = 0.0, q1 = 0.0;
        for (int c0 = lstart; c0 <= lmax; c0 += CHUNK) {
            if (c0 <= linj && ls == c0) {
                const double4 sd = seed[(int64_t)m * npairs + p];
                p2 = sd.x; p1 = sd.y; q2 = sd.z; q1 = sd.w;
            }
#pragma unroll
            for (int i = 0; i < CHUNK; ++i) {
                const int l = c0 + i;
                const double t = alpha[ro + l] * x;
                const double bl = beta[ro + l];
                const double na = fma(t + bl, p1, -p2);
                const double nb = fma(t - bl, q1, -q2);
                p2 = p1; p1 = na; q2 = q1; q1 = nb;
                const T a = (T)na, bt = (T)nb;
                const c2<T> u = ru[l], v = rv[l];
                if (i & 1) {
                    auB.x = fma(a, u.x, auB.x);  auB.y = fma(a, u.y, auB.y);
                    bvB.x = fma(bt, v.x, bvB.x); bvB.y = fma(bt, v.y, bvB.y);
                    buB.x = fma(bt, u.x, buB.x); buB.y = fma(bt, u.y, buB.y);
                    avB.x = fma(a, v.x, avB.x);  avB.y = fma(a, v.y, avB.y);
                } else {
                    auA.x = fma(a, u.x, auA.x);  auA.y = fma(a, u.y, auA.y);
                    bvA.x = fma(bt, v.x, bvA.x); bvA.y = fma(bt, v.y, bvA.y);
                    buA.x = fma(bt, u.x, buA.x); buA.y = fma(bt, u.y, buA.y);
                    avA.x = fma(a, v.x, avA.x);  avA.y = fma(a, v.y, avA.y);
                }
            }
        }
    }
    if (!valid) return;
    const T sigma = (s & 1) ? (T)-1 : (T)1;
    // (-1)^{l+m} = +1 on class A when par0 = 0
    const T sg = par0 ? (T)-1 : (T)1;
    // north: P = -sum a u, M = -sigma sum b v; south: P = -sum (-1)^{l+m} b u, M = -sigma sum (-1)^{l+m} a v
    const c2<T> Pn = mk<T>(-(auA.x + auB.x), -(auA.y + auB.y));
    const c2<T> Mn = mk<T>(-sigma * (bvA.x + bvB.x), -sigma * (bvA.y + bvB.y));
    const c2<T> Ps = mk<T>(-sg * (buA.x - buB.x), -sg * (buA.y - buB.y));
    const c2<T> Ms = mk<T>(-sigma * sg * (avA.x - avB.x), -sigma * sg * (avA.y - avB.y));
    // G_Q = (P + M)/2, G_U = -i (P - M)/2
    const c2<T> gqn = mk<T>((T)0.5 * (Pn.x + Mn.x), (T)0.5 * (Pn.y + Mn.y));
    const c2<T> gun = mk<T>((T)0.5 * (Pn.y - Mn.y), (T)-0.5 * (Pn.x - Mn.x));
    const c2<T> gqs = mk<T>((T)0.5 * (Ps.x + Ms.x), (T)0.5 * (Ps.y + Ms.y));
    const c2<T> gus = mk<T>((T)0.5 * (Ps.y - Ms.y), (T)-0.5 * (Ps.x - Ms.x));
    // synthesis layout: ring pair p, north and south halves
    c2<T> *oq = phase + (int64_t)(2 * b) * pstride + syn_ofs(m, 0, p, npairs);
    c2<T> *ou = oq + pstride;
    const int64_t south = (int64_t)pairs8(npairs) * MB;
    oq[0] = gqn;
    ou[0] = gun;
    if (nr - 1 - p != p) {
        oq[south] = gqs;
        ou[south] = gus;
    }
}

// ---------------------------------------------------------------------------
// synthesis on the matrix cores (FP32 plans).  Per m the l-sums of
// spin_synthesis_kernel are four products over l,
//   X_{f,c}[n][p] = sum_{l in class c} Y[n][l] f_l(p),   f = a, b;
//   c = e (l - l0 even), o (odd);  Y[n][l] = (u.re, u.im, v.re, v.im)[n%4]
// of polarisation pair n/4 of a group of 4 (u, v: the scaled rows of
// spin_prescale_kernel), and with sg = (-1)^{l0+m}, sigma = (-1)^s
//   P_n = -(X_ae + X_ao)[u],        M_n = -sigma (X_be + X_bo)[v],
//   P_s = -sg (X_be - X_bo)[u],     M_s = -sigma sg (X_ae - X_ao)[v],
//   G_Q = (P + M)/2,  G_U = -i (P - M)/2.
// On v_mfma_f32_16x16x4f32 with A = Y (16 rows n, 4 l of one class), B =
// lambda (4 l, 16 ring pairs): every wave works alone on 32 ring pairs of
// one m.  Per 16-l chunk lanes 0-31 run the FP64 a-recursion and lanes 32-63
// the b-recursion of the wave's pairs into the wave's own LDS tiles
// [f][l][pair]; after a wave barrier the wave contracts them (2 column tiles
// x 2 functions x 2 classes, 2 l-steps of 8).  Small per-wave tiles keep 5
// waves per SIMD resident, so one wave's recursion overlaps another's
// products.  The output lane holds the
// four columns of one polarisation pair for one ring pair: D lane =
// col + 16 (row/4), reg = row%4, so it combines them without exchanges.
// Lane layout of the 16x16x4 f32 instruction (checked on gfx950):
//   A: lane = row + 16 k   B: lane = col + 16 k   D: lane = col + 16 (row/4), reg = row%4
// ---------------------------------------------------------------------------
constexpr int SS_WAVES = 4;               // waves per workgroup (same m), 32 ring pairs each
constexpr int SS_L = 16;                  // l per chunk
constexpr int SS_ROW = 40;                // floats per tile row (32 pairs + 8): 2 rows = 16 mod 64 banks
constexpr int SS_TILE = SS_L * SS_ROW + 32;   // floats per function tile (b tile 32 banks from a's)
constexpr int SS_WAVE_LDS = 2 * SS_TILE;      // floats per wave

__global__ void __launch_bounds__(SS_WAVES * 64)
spin_synthesis_mfma_kernel(const float2 *__restrict__ scaled, int64_t nrow, float2 *__restrict__ phase,
                           int64_t pstride, const double *__restrict__ px, const int64_t *__restrict__ rofs,
                           const double *__restrict__ alpha, const double *__restrict__ beta,
                           const int32_t *__restrict__ lseed, const double4 *__restrict__ seed, int npairs, int nr,
                           int lmax, int mmax, int nsb, int npol, int s) {
    __shared__ float tiles[SS_WAVES * SS_WAVE_LDS];
    __shared__ double coefs[SS_WAVES][2][SS_L];     // per wave: alpha, beta of the chunk
    int mb, item;
    xcd_map(blockIdx.x, MB * nsb, &mb, &item);
    const int m = mb * MB + item / nsb;
    const int sb = item % nsb;
    if (m > mmax) return;
    const int g = blockIdx.y;                       // group of 4 polarisation pairs
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int r = lane & 31, h = lane >> 5;         // recursion: pair r of the wave, function h (0: a, 1: b)
    float *tw = tiles + w * SS_WAVE_LDS;

    const int pbase = (sb * SS_WAVES + w) * 32;
    const int p = pbase + r;
    double x = 0.0, u1 = 0.0, u2 = 0.0;
    double2 sd = make_double2(0.0, 0.0);
    int ls = SNEVER;
    if (p < npairs) {
        x = px[p];
        ls = lseed[(int64_t)m * npairs + p];
        if (ls <= lmax) {
            const double4 q = seed[(int64_t)m * npairs + p];
            sd = h ? make_double2(q.z, q.w) : make_double2(q.x, q.y);
        }
    }
    const int lstart = swmin(ls);
    const int winj = swmax(ls <= lmax ? ls : -1);

    const int l0 = max(m, abs(s));
    const int64_t ro = rofs[m] - m;                 // table / scaled-row index of l
    // A operand of this lane: Y[n = lane%16][l], k = lane/16
    const int n = lane & 15, k = lane >> 4;
    const int gb = g * 4 + (n >> 2), comp = n & 3;
    const float *yrow = reinterpret_cast<const float *>(scaled + (int64_t)(2 * min(gb, npol - 1) + (comp >> 1)) * nrow + ro) +
                        (comp & 1);
    const bool yok = gb < npol;
    float *lw = tw + h * SS_TILE + r;               // recursion writes: [h][l][r]
    const float *lr = tw + k * 2 * SS_ROW + n;      // B reads: [f][l = 8 st + 2 k + c][16 t + col]

    sfloatx4 D[2][2][2];                            // [column tile][function][class]
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int f = 0; f < 2; ++f)
#pragma unroll
            for (int c = 0; c < 2; ++c) D[t][f][c] = sfloatx4{0.f, 0.f, 0.f, 0.f};
    // chunk operands, loaded a chunk ahead: Y of this lane (SS_L/4 l) and the
    // chunk's alpha, beta (lanes i and i + 32 hold l = c0 + i), which the
    // wave stages in LDS for broadcast reads
    float yv[SS_L / 4];
    double av = 0.0, bv = 0.0;
    auto load = [&](int c0) {
#pragma unroll
        for (int st = 0; st < SS_L / 8; ++st)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int l = c0 + 8 * st + 2 * k + c;
                yv[2 * st + c] = (yok && l <= lmax) ? yrow[2 * (int64_t)l] : 0.0f;
            }
        av = alpha[ro + c0 + (r & (SS_L - 1))];
        bv = beta[ro + c0 + (r & (SS_L - 1))];
    };
    const double sgn = h ? -1.0 : 1.0;              // b: beta -> -beta
    const double *ca = coefs[w][0], *cb = coefs[w][1];
    if (lstart <= lmax) load(lstart);
    for (int c0 = lstart; c0 <= lmax; c0 += SS_L) {
        float yc[SS_L / 4];
#pragma unroll
        for (int i = 0; i < SS_L / 4; ++i) yc[i] = yv[i];
        if (r < SS_L) coefs[w][h][r] = h ? bv : av;
        if (c0 + SS_L <= lmax) load(c0 + SS_L);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // 1. recursion of this lane's (pair, function) over l = c0 .. c0+SS_L-1;
        // past lmax the tables run into the next row: those values stay out
        // of the tile (Y is zero there, but 0 * inf would not be)
        const bool tail = c0 + SS_L > lmax + 1;
#pragma unroll
        for (int tt = 0; tt < SS_L; ++tt) {
            if ((tt & 7) == 0 && c0 <= winj && ls == c0 + tt) {
                u2 = sd.x;
                u1 = sd.y;
            }
            const double t = ca[tt] * x;
            const double u = fma(fma(sgn, cb[tt], t), u1, -u2);
            u2 = u1;
            u1 = u;
            lw[tt * SS_ROW] = (!tail || c0 + tt <= lmax) ? (float)u : 0.0f;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // 2. the wave's products over the chunk: 4 steps of 8 l
#pragma unroll
        for (int st = 0; st < SS_L / 8; ++st) {
#pragma unroll
            for (int c = 0; c < 2; ++c) {
#pragma unroll
                for (int f = 0; f < 2; ++f)
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        const float lam = lr[f * SS_TILE + (8 * st + c) * SS_ROW + 16 * t];
                        D[t][f][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(yc[2 * st + c], lam, D[t][f][c], 0, 0, 0);
                    }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    // this lane: ring pair pbase + 16 t + (lane & 15), polarisation pair
    // g*4 + lane/16, regs = (u.re, u.im, v.re, v.im) columns
    const int jo = g * 4 + (lane >> 4);
    if (jo >= npol) return;
    const float sigma = (s & 1) ? -1.0f : 1.0f;
    const float sg = ((l0 + m) & 1) ? -1.0f : 1.0f;
    // synthesis layout: ring pair pp, north and south halves
    float2 *oq = phase + (int64_t)(2 * jo) * pstride + syn_ofs(m, 0, 0, npairs);
    float2 *ou = oq + pstride;
    const int64_t south = (int64_t)pairs8(npairs) * MB;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int pp = pbase + 16 * t + (lane & 15);
        if (pp >= npairs) continue;
        const sfloatx4 ae = D[t][0][0], ao = D[t][0][1], be = D[t][1][0], bo = D[t][1][1];
        const float pnr = -(ae[0] + ao[0]), pni = -(ae[1] + ao[1]);
        const float mnr = -sigma * (be[2] + bo[2]), mni = -sigma * (be[3] + bo[3]);
        const float psr = -sg * (be[0] - bo[0]), psi = -sg * (be[1] - bo[1]);
        const float msr = -sigma * sg * (ae[2] - ao[2]), msi = -sigma * sg * (ae[3] - ao[3]);
        const int64_t o = (pp >> 3) * (8 * MB) + (pp & 7);
        // G_Q = (P + M)/2, G_U = -i (P - M)/2
        oq[o] = make_float2(0.5f * (pnr + mnr), 0.5f * (pni + mni));
        ou[o] = make_float2(0.5f * (pni - mni), -0.5f * (pnr - mnr));
        if (nr - 1 - pp != pp) {
            oq[south + o] = make_float2(0.5f * (psr + msr), 0.5f * (psi + msi));
            ou[south + o] = make_float2(0.5f * (psi - msi), -0.5f * (psr - msr));
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// ring-pair slots per lane: the plan's tile, else up to 4 (more costs
// occupancy: ~56 VGPRs a slot)
int spin_slots_of(const Plan *p) {
    if (p->tile_spin_slots > 0) return p->tile_spin_slots;
    int r = 1;
    while (r < 4 && (int64_t)r * 64 < p->npairs) r *= 2;
    return r;
}

int spin_groups(const Plan *p) {
    const int r = spin_slots_of(p);
    return (int)((p->npairs + 64 * r - 1) / (64 * r));
}

template <typename T>
static int spin_analysis_t(const Plan *p, int nmaps, const c2<T> *phase, c2<T> *alm, int64_t astride, c2<T> *partial,
                           int adjoint, hipStream_t s) {
    const int R = spin_slots_of(p);
    const int G = spin_groups(p);
    const int nmb8 = (p->nmb + 7) / 8 * 8;
    dim3 grid(nmb8 * MB * G, nmaps);
    const int64_t pstride = phase_elems_per_map(p);
#define HBX_SA(RV)                                                                                            \
    hipLaunchKernelGGL((spin_analysis_kernel<T, RV>), grid, dim3(64), 0, s, phase, pstride, alm, astride, partial, \
                       p->d_px, p->d_rofs, p->d_alpha, p->d_beta, p->d_sl, p->d_lseed, p->d_seed4, p->npairs,      \
                       p->nrings, p->lmax, p->mmax, G, p->nlm, p->spin, adjoint)
    timer_mark(HBX_TIMER_SPIN_ANALYSIS, true, s);
    switch (R) {
    case 1: HBX_SA(1); break;
    case 2: HBX_SA(2); break;
    case 4: HBX_SA(4); break;
    default: return set_error(HBX_EUNSUPPORTED, "spin slot count");
    }
#undef HBX_SA
    timer_mark(HBX_TIMER_SPIN_ANALYSIS, false, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "spin_analysis_kernel launch");
    if (G > 1) return launch_group_reduce(p, 2 * nmaps, partial, alm, astride, G, p->nlm, 0, p->mmax,
                                          sizeof(T) == 8, adjoint, s);
    return 0;
}

// FP64 band of FP32 spin plans: E/B rows m < M_LO from the FP64 band phase
// of Q and U (ring_band_kernel), with the FP64 kernel, then rounded into alm
int launch_spin_analysis_band(const Plan *p, int npol, const void *band_phase, void *alm, int64_t astride,
                              void *band_alm, void *band_partial, int adjoint, hipStream_t s) {
    const int R = spin_slots_of(p);
    const int G = spin_groups(p);
    const int mtop = std::min(M_LO - 1, p->mmax);
    const int64_t nb = band_nlm(p);
    dim3 grid(8 * MB * G, npol);                   // xcd_map over 8 m-blocks; only m-block 0 is work
#define HBX_SB(RV)                                                                                            \
    hipLaunchKernelGGL((spin_analysis_kernel<double, RV>), grid, dim3(64), 0, s, (const double2 *)band_phase,   \
                       band_phase_elems(p), (double2 *)band_alm, nb, (double2 *)band_partial, p->d_px,         \
                       p->d_rofs, p->d_alpha, p->d_beta, p->d_sl, p->d_lseed, p->d_seed4, p->npairs, p->nrings, \
                       p->lmax, mtop, G, nb, p->spin, adjoint)
    switch (R) {
    case 1: HBX_SB(1); break;
    case 2: HBX_SB(2); break;
    case 4: HBX_SB(4); break;
    default: return set_error(HBX_EUNSUPPORTED, "spin slot count");
    }
#undef HBX_SB
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "spin_analysis_kernel (band) launch");
    int rc;
    if (G > 1 && (rc = launch_group_reduce(p, 2 * npol, band_partial, band_alm, nb, G, nb, 0, mtop, true, adjoint, s)))
        return rc;
    return launch_band_to_f32(p, 2 * npol, band_alm, alm, astride, s);
}

int launch_spin_analysis(const Plan *p, int nmaps, const void *phase, void *alm, int64_t astride, void *partial,
                         int adjoint, hipStream_t s) {
    if (p->precision == HBX_F64)
        return spin_analysis_t<double>(p, nmaps, (const double2 *)phase, (double2 *)alm, astride, (double2 *)partial,
                                       adjoint, s);
    return spin_analysis_t<float>(p, nmaps, (const float2 *)phase, (float2 *)alm, astride, (float2 *)partial,
                                  adjoint, s);
}

// FP32 batches of at least mfma_min_maps maps (Q and U count) contract on the
// matrix cores; FP64 plans and smaller batches run the VALU kernels
bool spin_uses_mfma(const Plan *p, int npol) { return p->precision == HBX_F32 && 2 * npol >= mfma_min_maps(p); }

int spin_mfma_groups(const Plan *p) { return (p->npairs + SpinMfGeom::P - 1) / SpinMfGeom::P; }

// rows m >= M_LO of FP32 plans (the FP64 band covers m < M_LO); lfirst holds
// (mmax + 1) x spin_mfma_groups entries
int launch_spin_analysis_mfma(const Plan *p, int npol, const void *phase, void *alm, int64_t astride, void *partial,
                              int32_t *lfirst, int adjoint, hipStream_t s) {
    if (p->mmax < M_LO) return 0;
    const int G = spin_mfma_groups(p);
    const int nmb8 = (p->nmb + 7) / 8 * 8;
    dim3 grid(nmb8 * MB * G, (npol + 3) / 4);
    timer_mark(HBX_TIMER_SPIN_ANALYSIS, true, s);
    hipLaunchKernelGGL(spin_analysis_mfma_kernel, grid, dim3(SpinMfGeom::W * 64), SpinMfGeom::LDS, s,
                       (const float2 *)phase, phase_elems_per_map(p), (float2 *)alm, astride, (float2 *)partial,
                       p->d_px, p->d_rofs, p->d_alpha, p->d_beta, p->d_sl, p->d_lseed, p->d_seed4, p->npairs,
                       p->nrings, p->lmax, p->mmax, G, p->nlm, npol, p->spin, adjoint, M_LO, lfirst);
    timer_mark(HBX_TIMER_SPIN_ANALYSIS, false, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "spin_analysis_mfma_kernel launch");
    if (G > 1)
        return launch_group_reduce(p, 2 * npol, partial, alm, astride, G, p->nlm, M_LO, p->mmax, false, adjoint, s,
                                   lfirst);
    return 0;
}

int launch_spin_prescale(const Plan *p, int npol, const void *alm, int64_t astride, void *scaled, int adjoint,
                         hipStream_t s) {
    dim3 grid(p->mmax + 1, npol);
    if (p->precision == HBX_F64)
        hipLaunchKernelGGL(spin_prescale_kernel<double>, grid, dim3(256), 0, s, (const double2 *)alm, astride,
                           (double2 *)scaled, p->nrow, p->d_rofs, p->d_sl, p->lmax, adjoint);
    else
        hipLaunchKernelGGL(spin_prescale_kernel<float>, grid, dim3(256), 0, s, (const float2 *)alm, astride,
                           (float2 *)scaled, p->nrow, p->d_rofs, p->d_sl, p->lmax, adjoint);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "spin_prescale_kernel launch");
    return 0;
}

int launch_spin_synthesis(const Plan *p, int nmaps, const void *scaled, void *phase, hipStream_t s) {
    const int nmb8 = (p->nmb + 7) / 8 * 8;
    if (spin_uses_mfma(p, nmaps)) {
        const int nsb = (p->npairs + SS_WAVES * 32 - 1) / (SS_WAVES * 32);
        dim3 grid(nmb8 * MB * nsb, (nmaps + 3) / 4);
        timer_mark(HBX_TIMER_SPIN_SYNTHESIS, true, s);
        hipLaunchKernelGGL(spin_synthesis_mfma_kernel, grid, dim3(SS_WAVES * 64), 0, s, (const float2 *)scaled,
                           p->nrow, (float2 *)phase, phase_elems_per_map(p), p->d_px, p->d_rofs, p->d_alpha,
                           p->d_beta, p->d_lseed, p->d_seed4, p->npairs, p->nrings, p->lmax, p->mmax, nsb, nmaps,
                           p->spin);
        timer_mark(HBX_TIMER_SPIN_SYNTHESIS, false, s);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_error(e, "spin_synthesis_mfma_kernel launch");
        return 0;
    }
    const int npb = (p->npairs + 63) / 64;
    dim3 grid(nmb8 * npb, nmaps);
    const int64_t pstride = phase_elems_per_map(p);
    timer_mark(HBX_TIMER_SPIN_SYNTHESIS, true, s);
    if (p->precision == HBX_F64)
        hipLaunchKernelGGL(spin_synthesis_kernel<double>, grid, dim3(512), 0, s, (const double2 *)scaled, p->nrow,
                           (double2 *)phase, pstride, p->d_px, p->d_rofs, p->d_alpha, p->d_beta, p->d_lseed,
                           p->d_seed4, p->npairs, p->nrings, p->lmax, p->mmax, p->nmb, npb, p->spin);
    else
        hipLaunchKernelGGL(spin_synthesis_kernel<float>, grid, dim3(512), 0, s, (const float2 *)scaled, p->nrow,
                           (float2 *)phase, pstride, p->d_px, p->d_rofs, p->d_alpha, p->d_beta, p->d_lseed,
                           p->d_seed4, p->npairs, p->nrings, p->lmax, p->mmax, p->nmb, npb, p->spin);
    timer_mark(HBX_TIMER_SPIN_SYNTHESIS, false, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "spin_synthesis_kernel launch");
    return 0;
}

}  // namespace hbx
