// Ring phase stage: batched real DFTs over HEALPix rings of uneven length.
//
// Reference: per-ring cupy.fft.fft(ring)/nphi (sht_raw.py:38-43,
// sht_fallback.py:51-56) and cupy.fft.ifft(F)*nphi (isht.py:24), with the
// alias rule of sht_raw.py:51 (ring contributes to m < nphi only).
//
// One workgroup transforms one ring of one map entirely in LDS:
//   * a real ring of length n = 2h is packed as h complex values
//     z_j = x_{2j} + i x_{2j+1} (one half-length complex DFT per ring);
//   * h a power of two  -> radix-4/2 Stockham FFT of size M = h in LDS;
//   * otherwise         -> Bluestein chirp-z with a power-of-two FFT of size
//     M >= 2h-1 and a per-length chirp spectrum precomputed at plan time;
//   * the half-length spectrum is untangled into X[k], k <= h, and written as
//     F_r(m) = scale_r * e^{-i m phi0_r} X[m mod n]  for m < mcut_r, else 0,
//     into the m-blocked ring-major phase buffer [map][m/8][ring][m%8].
// Rings are grouped into size classes (one launch per power-of-two M) so
// that each launch reserves exactly the LDS it needs.
#include <algorithm>
#include <map>

#include "hbx_device.h"
#include "hbx_internal.h"

namespace hbx {

// ---------------------------------------------------------------------------
// Twiddles w^J = e^{-2 pi i J / 2^twshift} come from a two-level FP64 table in
// LDS (staged from the plan's FP64 global table once per workgroup):
//   w^J = coarse[J >> f] * fine[J & (2^f - 1)],  f = min(7, twshift),
// 2^(twshift-f) + 2^f entries (192 for 8192-point rings).  The product is
// formed in FP64 and rounded once, so FP32 transforms get correctly rounded
// twiddles (the FP32 ring stage's error budget is set by them), and the FFT
// passes read LDS instead of waiting on global loads between barriers.
// ---------------------------------------------------------------------------
constexpr int TW_FINE_LOG = 7;
inline int tw_fine_log(int twshift) { return twshift < TW_FINE_LOG ? twshift : TW_FINE_LOG; }
__device__ __forceinline__ int tw_fine_log_d(int twshift) { return min(twshift, TW_FINE_LOG); }
inline int tw_lds_entries(int twshift) {
    const int f = tw_fine_log(twshift);
    return (1 << (twshift - f)) + (1 << f);
}
// LDS bytes of a ring workgroup: M data elements + the FP64 twiddle table
template <typename T> inline size_t ring_lds_bytes(int M, int twshift) {
    return (size_t)M * sizeof(c2<T>) + (size_t)tw_lds_entries(twshift) * sizeof(double2);
}

__device__ void stage_twiddles(double2 *tab, const double2 *__restrict__ tw, int twshift) {
    const int f = tw_fine_log_d(twshift), nc = 1 << (twshift - f), nf = 1 << f;
    for (int i = threadIdx.x; i < nc + nf; i += blockDim.x) tab[i] = (i < nc) ? tw[i << f] : tw[i - nc];
}

template <typename T>
__device__ __forceinline__ c2<T> tw_get(const double2 *tab, int J, int f, int nc) {
    const double2 a = tab[J >> f], b = tab[nc + (J & ((1 << f) - 1))];
    return mk<T>((T)(a.x * b.x - a.y * b.y), (T)(a.x * b.y + a.y * b.x));
}

// ---------------------------------------------------------------------------
// in-LDS FFT of size M (power of two), forward (e^{-i}) or inverse (e^{+i},
// unnormalised).  Stockham radix-4 passes plus one radix-2 pass when log2 M
// is odd; each pass stages its inputs in registers so one buffer suffices.
// tab: the LDS twiddle table (stage_twiddles, visible to every thread).
// ---------------------------------------------------------------------------
template <typename T, int NB>
__device__ void lds_fft(c2<T> *buf, int M, int log2M, const double2 *tab, int twshift, bool inv) {
    const int nt = blockDim.x, tid = threadIdx.x;
    const int tf = tw_fine_log_d(twshift), tnc = 1 << (twshift - tf);
    const int Q = M >> 2;
    int p = 1, lp = 0;
    const int n4 = log2M >> 1;
    for (int s = 0; s < n4; ++s) {
        c2<T> v[NB][4];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int i = tid + b * nt;
            if (i < Q) {
#pragma unroll
                for (int j = 0; j < 4; ++j) v[b][j] = buf[i + j * Q];
            }
        }
        __syncthreads();
        const int sh = twshift - lp - 2;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int i = tid + b * nt;
            if (i < Q) {
                const int k = i & (p - 1);
                c2<T> w1 = tw_get<T>(tab, k << sh, tf, tnc), w2 = tw_get<T>(tab, (2 * k) << sh, tf, tnc),
                      w3 = tw_get<T>(tab, (3 * k) << sh, tf, tnc);
                if (inv) { w1.y = -w1.y; w2.y = -w2.y; w3.y = -w3.y; }
                const c2<T> x0 = v[b][0];
                const c2<T> x1 = cmul<T>(v[b][1], w1);
                const c2<T> x2 = cmul<T>(v[b][2], w2);
                const c2<T> x3 = cmul<T>(v[b][3], w3);
                const c2<T> s02 = cadd<T>(x0, x2), d02 = csub<T>(x0, x2);
                const c2<T> s13 = cadd<T>(x1, x3), d13 = csub<T>(x1, x3);
                const c2<T> r13 = inv ? cmul_pi<T>(d13) : cmul_mi<T>(d13);
                const int j0 = ((i - k) << 2) + k;
                buf[j0] = cadd<T>(s02, s13);
                buf[j0 + p] = cadd<T>(d02, r13);
                buf[j0 + 2 * p] = csub<T>(s02, s13);
                buf[j0 + 3 * p] = csub<T>(d02, r13);
            }
        }
        __syncthreads();
        p <<= 2;
        lp += 2;
    }
    if (log2M & 1) {
        const int H = M >> 1;  // p == H here
        const int sh = twshift - log2M;
#pragma unroll
        for (int b = 0; b < 2 * NB; ++b) {
            const int i = tid + b * nt;
            if (i < H) {
                c2<T> w = tw_get<T>(tab, i << sh, tf, tnc);
                if (inv) w.y = -w.y;
                const c2<T> x0 = buf[i];
                const c2<T> x1 = cmul<T>(buf[i + H], w);
                buf[i] = cadd<T>(x0, x1);
                buf[i + H] = csub<T>(x0, x1);
            }
        }
        __syncthreads();
    }
}

// chirp c_j = exp(i pi j^2 / h), with j^2 reduced mod 2h exactly
template <typename T>
__device__ __forceinline__ c2<T> chirp(int j, int h) {
    const long long q = ((long long)j * j) % (2LL * h);
    return expipi<T>((T)q / (T)h);
}

template <typename T> __device__ __forceinline__ T load_as(const double *p) { return (T)(*p); }
template <typename T> __device__ __forceinline__ T load_as(const float *p) { return (T)(*p); }

// ---------------------------------------------------------------------------
// analysis: maps -> phase
// ---------------------------------------------------------------------------
template <typename T, typename TIn, int NB>
__global__ void __launch_bounds__(1024)
ring_analysis_kernel(const TIn *__restrict__ maps, int64_t map_stride, c2<T> *__restrict__ phase,
                     int64_t phase_stride, const int32_t *__restrict__ rings,
                     const int64_t *__restrict__ ring_start, const int32_t *__restrict__ nphi,
                     const int32_t *__restrict__ mcut, const double *__restrict__ phi0,
                     const double *__restrict__ scale, const double2 *__restrict__ tw, int twshift,
                     const c2<T> *__restrict__ V, const int64_t *__restrict__ vofs, int M, int log2M,
                     int mmax, int nr) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    c2<T> *buf = reinterpret_cast<c2<T> *>(smem_raw);
    double2 *tab = reinterpret_cast<double2 *>(buf + M);
    const int r = rings[blockIdx.x];
    const int b = blockIdx.y;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int n = nphi[r], h = n >> 1;
    const bool pow2 = (h == M);
    const TIn *f = maps + (int64_t)b * map_stride + ring_start[r];
    stage_twiddles(tab, tw, twshift);

    for (int j = tid; j < M; j += nt) {
        c2<T> z = mk<T>(0, 0);
        if (j < h) {
            z = mk<T>(load_as<T>(f + 2 * j), load_as<T>(f + 2 * j + 1));
            if (!pow2) z = cmul<T>(z, cconj<T>(chirp<T>(j, h)));
        }
        buf[j] = z;
    }
    __syncthreads();
    lds_fft<T, NB>(buf, M, log2M, tab, twshift, false);
    if (!pow2) {
        const c2<T> *Vr = V + vofs[r];
        for (int k = tid; k < M; k += nt) buf[k] = cmul<T>(buf[k], Vr[k]);
        __syncthreads();
        lds_fft<T, NB>(buf, M, log2M, tab, twshift, true);
        const T invM = (T)1 / (T)M;
        for (int k = tid; k < h; k += nt)
            buf[k] = cscale<T>(cmul<T>(buf[k], cconj<T>(chirp<T>(k, h))), invM);
        __syncthreads();
    }
    const T sc = (T)(scale ? scale[r] : 1.0);
    const int cut = mcut[r];
    const double ph0 = phi0 ? phi0[r] : 0.0;
    c2<T> *out = phase + (int64_t)b * phase_stride;
    for (int m = tid; m <= mmax; m += nt) {
        c2<T> F = mk<T>(0, 0);
        if (m < cut) {
            int k = m % n;
            bool cj = false;
            if (k > h) { k = n - k; cj = true; }
            const int ia = (k == h) ? 0 : k;
            const int ib = (k == 0) ? 0 : h - k;
            const c2<T> za = buf[ia], zb = cconj<T>(buf[ib]);
            const c2<T> A = cscale<T>(cadd<T>(za, zb), (T)0.5);
            const c2<T> D = cscale<T>(csub<T>(za, zb), (T)0.5);
            const c2<T> W = expipi<T>((T)(-2.0 * (double)k / (double)n));
            c2<T> X = cadd<T>(A, cmul<T>(W, cmul_mi<T>(D)));
            if (cj) X = cconj<T>(X);
            F = cscale<T>(X, sc);
            if (ph0 != 0.0) {
                double s, c;
                sincos(-(double)m * ph0, &s, &c);
                F = cmul<T>(F, mk<T>((T)c, (T)s));
            }
        }
        out[((int64_t)(m >> 3) * nr + r) * MB + (m & 7)] = F;
    }
}

// ---------------------------------------------------------------------------
// synthesis: phase -> maps
// ---------------------------------------------------------------------------
template <typename T, typename TOut, int NB>
__global__ void __launch_bounds__(1024)
ring_synthesis_kernel(const c2<T> *__restrict__ phase, int64_t phase_stride, TOut *__restrict__ maps,
                      int64_t map_stride, const int32_t *__restrict__ rings,
                      const int64_t *__restrict__ ring_start, const int32_t *__restrict__ nphi,
                      const int32_t *__restrict__ mcut, const double *__restrict__ phi0,
                      const double *__restrict__ scale, const double2 *__restrict__ tw, int twshift,
                      const c2<T> *__restrict__ V, const int64_t *__restrict__ vofs, int M, int log2M,
                      int mmax, int nr) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    c2<T> *buf = reinterpret_cast<c2<T> *>(smem_raw);
    double2 *tab = reinterpret_cast<double2 *>(buf + M);
    const int r = rings[blockIdx.x];
    const int b = blockIdx.y;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int n = nphi[r], h = n >> 1;
    const bool pow2 = (h == M);
    const int top = min(mmax, mcut[r] - 1);
    stage_twiddles(tab, tw, twshift);
    const double ph0 = phi0 ? phi0[r] : 0.0;
    const c2<T> *G = phase + (int64_t)b * phase_stride;

    auto gval = [&](int m) -> c2<T> {
        c2<T> g = G[((int64_t)(m >> 3) * nr + r) * MB + (m & 7)];
        if (m == 0) g.y = 0;
        if (ph0 != 0.0) {
            double s, c;
            sincos((double)m * ph0, &s, &c);
            g = cmul<T>(g, mk<T>((T)c, (T)s));
        }
        return g;
    };
    // X[k] = sum_{m = k mod n} G'(m) + sum_{m = -k mod n, m >= 1} conj G'(m)
    auto xval = [&](int k) -> c2<T> {
        c2<T> acc = mk<T>(0, 0);
        for (int m = k; m <= top; m += n) acc = cadd<T>(acc, gval(m));
        for (int m = (k == 0 ? n : n - k); m <= top; m += n) acc = cadd<T>(acc, cconj<T>(gval(m)));
        return acc;
    };
    for (int k = tid; k <= (h >> 1); k += nt) {
        const c2<T> Xa = xval(k), Xb = xval(h - k);
        {
            const c2<T> s = cadd<T>(Xa, cconj<T>(Xb)), d = csub<T>(Xa, cconj<T>(Xb));
            const c2<T> W = expipi<T>((T)(2.0 * (double)k / (double)n));
            c2<T> z = cadd<T>(s, cmul_pi<T>(cmul<T>(W, d)));
            if (!pow2) z = cconj<T>(cmul<T>(z, chirp<T>(k, h)));
            buf[k] = z;
        }
        const int k2 = h - k;
        if (k != 0 && k2 != k) {
            const c2<T> s = cadd<T>(Xb, cconj<T>(Xa)), d = csub<T>(Xb, cconj<T>(Xa));
            const c2<T> W = expipi<T>((T)(2.0 * (double)k2 / (double)n));
            c2<T> z = cadd<T>(s, cmul_pi<T>(cmul<T>(W, d)));
            if (!pow2) z = cconj<T>(cmul<T>(z, chirp<T>(k2, h)));
            buf[k2] = z;
        }
    }
    if (!pow2)
        for (int j = h + tid; j < M; j += nt) buf[j] = mk<T>(0, 0);
    __syncthreads();
    if (pow2) {
        lds_fft<T, NB>(buf, M, log2M, tab, twshift, true);
    } else {
        lds_fft<T, NB>(buf, M, log2M, tab, twshift, false);
        const c2<T> *Vr = V + vofs[r];
        for (int k = tid; k < M; k += nt) buf[k] = cmul<T>(buf[k], Vr[k]);
        __syncthreads();
        lds_fft<T, NB>(buf, M, log2M, tab, twshift, true);
        const T invM = (T)1 / (T)M;
        for (int j = tid; j < h; j += nt)
            buf[j] = cscale<T>(cmul<T>(chirp<T>(j, h), cconj<T>(buf[j])), invM);
        __syncthreads();
    }
    const T sc = (T)(scale ? scale[r] : 1.0);
    TOut *f = maps + (int64_t)b * map_stride + ring_start[r];
    for (int j = tid; j < h; j += nt) {
        const c2<T> z = buf[j];
        f[2 * j] = (TOut)(z.x * sc);
        f[2 * j + 1] = (TOut)(z.y * sc);
    }
}

// ---------------------------------------------------------------------------
// FP64 low-m band of FP32 plans: direct DFT in FP64 for m < M_LO,
//   F_r(m) = [m < mcut_r] scale_r e^{-i m phi0_r} sum_j f_j e^{-2 pi i m j / n}
// (same conventions as ring_analysis_kernel), one workgroup per (ring, map).
// e^{-2 pi i j/n} comes from sincospi per pixel, its powers by recurrence.
// ---------------------------------------------------------------------------
template <typename TIn>
__global__ void __launch_bounds__(256)
ring_band_kernel(const TIn *__restrict__ maps, int64_t map_stride, double2 *__restrict__ band, int64_t band_stride,
                 const int64_t *__restrict__ ring_start, const int32_t *__restrict__ nphi,
                 const int32_t *__restrict__ mcut, const double *__restrict__ phi0, const double *__restrict__ scale,
                 int mtop, int nr) {
    __shared__ double part[4][2 * M_LO];
    const int r = blockIdx.x, b = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int n = nphi[r];
    const TIn *f = maps + (int64_t)b * map_stride + ring_start[r];
    double re[M_LO], im[M_LO];
#pragma unroll
    for (int m = 0; m < M_LO; ++m) { re[m] = 0.0; im[m] = 0.0; }
    const double inv = 2.0 / (double)n;
    for (int j = tid; j < n; j += blockDim.x) {
        const double v = (double)f[j];
        double sn, cs;
        sincospi(-(double)j * inv, &sn, &cs);
        double wr = 1.0, wi = 0.0;
#pragma unroll
        for (int m = 0; m < M_LO; ++m) {
            re[m] = fma(v, wr, re[m]);
            im[m] = fma(v, wi, im[m]);
            const double t = wr * cs - wi * sn;
            wi = wr * sn + wi * cs;
            wr = t;
        }
    }
#pragma unroll
    for (int m = 0; m < M_LO; ++m) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            re[m] += __shfl_xor(re[m], o, 64);
            im[m] += __shfl_xor(im[m], o, 64);
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int m = 0; m < M_LO; ++m) { part[w][2 * m] = re[m]; part[w][2 * m + 1] = im[m]; }
    }
    __syncthreads();
    if (tid <= mtop) {
        const int m = tid;
        double x = 0.0, y = 0.0;
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) { x += part[q][2 * m]; y += part[q][2 * m + 1]; }
        double2 F = make_double2(0.0, 0.0);
        if (m < mcut[r]) {
            const double sc = scale ? scale[r] : 1.0;
            x *= sc;
            y *= sc;
            const double ph = phi0 ? phi0[r] : 0.0;
            if (ph != 0.0) {
                double sn, cs;
                sincos(-(double)m * ph, &sn, &cs);
                F = make_double2(x * cs - y * sn, x * sn + y * cs);
            } else {
                F = make_double2(x, y);
            }
        }
        band[(int64_t)b * band_stride + ((int64_t)(m >> 3) * nr + r) * MB + (m & 7)] = F;
    }
}

int launch_ring_band(const Plan *p, int nmaps, const void *maps, int map_dtype, int64_t map_stride, void *band,
                     const double *scale, hipStream_t s) {
    const int mtop = std::min(M_LO - 1, p->mmax);
    dim3 grid(p->nrings, nmaps);
    if (map_dtype == HBX_DT_F64)
        hipLaunchKernelGGL(ring_band_kernel<double>, grid, dim3(256), 0, s, (const double *)maps, map_stride,
                           (double2 *)band, band_phase_elems(p), p->d_ring_start, p->d_nphi, p->d_mcut, p->d_phi0,
                           scale, mtop, p->nrings);
    else
        hipLaunchKernelGGL(ring_band_kernel<float>, grid, dim3(256), 0, s, (const float *)maps, map_stride,
                           (double2 *)band, band_phase_elems(p), p->d_ring_start, p->d_nphi, p->d_mcut, p->d_phi0,
                           scale, mtop, p->nrings);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "ring_band_kernel launch");
    return 0;
}

// ---------------------------------------------------------------------------
// plan-time tables
// ---------------------------------------------------------------------------
template <typename T>
__global__ void twiddle_kernel(c2<T> *tw, int twn) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < twn) {
        double s, c;
        sincospi(-2.0 * (double)k / (double)twn, &s, &c);
        tw[k] = mk<T>((T)c, (T)s);
    }
}

// V_h = FFT_M(v), v_j = c_j (0 <= j < h), v_{M-j} = c_j (1 <= j < h), else 0
template <typename T, int NB>
__global__ void __launch_bounds__(1024)
chirp_spectrum_kernel(const int32_t *__restrict__ hs, const int64_t *__restrict__ ofs, c2<T> *V,
                      const double2 *__restrict__ tw, int twshift, int M, int log2M) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    c2<T> *buf = reinterpret_cast<c2<T> *>(smem_raw);
    double2 *tab = reinterpret_cast<double2 *>(buf + M);
    const int h = hs[blockIdx.x];
    stage_twiddles(tab, tw, twshift);
    for (int j = threadIdx.x; j < M; j += blockDim.x) {
        c2<T> v = mk<T>(0, 0);
        if (j < h) v = chirp<T>(j, h);
        else if (j > M - h) v = chirp<T>(M - j, h);
        buf[j] = v;
    }
    __syncthreads();
    lds_fft<T, NB>(buf, M, log2M, tab, twshift, false);
    c2<T> *out = V + ofs[blockIdx.x];
    for (int j = threadIdx.x; j < M; j += blockDim.x) out[j] = buf[j];
}

// threads per ring workgroup: 2 radix-4 butterflies per thread for M >= 4096
// (8192-point Bluestein rings need 1024 threads to stay at ~60 VGPRs, so
// two workgroups share a CU)
static int class_threads(int M) { return M >= 8192 ? 1024 : (M >= 4096 ? 512 : 256); }
static int class_nb(int M) {
    const int nt = class_threads(M);
    return std::max(1, M / (4 * nt));
}

// dispatch on NB for a kernel template taking <T, ..., NB>
#define HBX_NB_SWITCH(nb, ...)                                                \
    switch (nb) {                                                             \
    case 1: { constexpr int NBV = 1; __VA_ARGS__; } break;                    \
    case 2: { constexpr int NBV = 2; __VA_ARGS__; } break;                    \
    case 4: { constexpr int NBV = 4; __VA_ARGS__; } break;                    \
    case 8: { constexpr int NBV = 8; __VA_ARGS__; } break;                    \
    default: return set_error(HBX_EUNSUPPORTED, "FFT size class too large");  \
    }

template <typename T, typename K>
static int allow_lds(K kernel, size_t bytes) {
    if (bytes > 65536) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return hip_error(e, "hipFuncSetAttribute(MaxDynamicSharedMemorySize)");
    }
    return 0;
}

template <typename T>
static int build_chirp_spectra(Plan *p, const std::vector<int32_t> &hs_host,
                               const std::vector<int64_t> &ofs_host, int M, c2<T> *V,
                               const double2 *tw) {
    if (hs_host.empty()) return 0;
    int32_t *d_hs = nullptr;
    int64_t *d_ofs = nullptr;
    HBX_HIP(hipMalloc(&d_hs, hs_host.size() * sizeof(int32_t)));
    HBX_HIP(hipMalloc(&d_ofs, ofs_host.size() * sizeof(int64_t)));
    HBX_HIP(hipMemcpy(d_hs, hs_host.data(), hs_host.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HBX_HIP(hipMemcpy(d_ofs, ofs_host.data(), ofs_host.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    const int nt = class_threads(M), nb = class_nb(M), lg = ilog2(M);
    const size_t lds = ring_lds_bytes<T>(M, p->twshift);
    int rc = 0;
    HBX_NB_SWITCH(nb, {
        auto k = chirp_spectrum_kernel<T, NBV>;
        rc = allow_lds<T>(k, lds);
        if (rc == 0) {
            hipLaunchKernelGGL(k, dim3((unsigned)hs_host.size()), dim3(nt), lds, 0, d_hs, d_ofs, V, tw,
                               p->twshift, M, lg);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) rc = hip_error(e, "chirp_spectrum_kernel");
        }
    });
    hipError_t e = hipDeviceSynchronize();
    (void)hipFree(d_hs);
    (void)hipFree(d_ofs);
    if (rc) return rc;
    if (e != hipSuccess) return hip_error(e, "chirp_spectrum_kernel sync");
    return 0;
}

int plan_build_fft_tables(Plan *p, const std::vector<int32_t> &nphi) {
    const int nr = p->nrings;
    // class per ring
    std::map<int, std::vector<int32_t>> cls;            // M -> rings
    std::map<int, int64_t> vofs_of_h;                   // h -> offset (non-pow2)
    std::map<int, std::vector<int32_t>> hs_by_M;        // M -> distinct h
    std::map<int, std::vector<int64_t>> ofs_by_M;
    std::vector<int64_t> vofs(nr, -1);
    int maxM = 4;
    int64_t vtotal = 0;
    for (int r = 0; r < nr; ++r) {
        const int h = nphi[r] / 2;
        const bool pow2 = (h & (h - 1)) == 0;
        const int M = pow2 ? h : (1 << ilog2(2 * (int64_t)h - 1));
        cls[M].push_back(r);
        maxM = std::max(maxM, M);
        if (!pow2) {
            auto it = vofs_of_h.find(h);
            if (it == vofs_of_h.end()) {
                vofs_of_h[h] = vtotal;
                hs_by_M[M].push_back(h);
                ofs_by_M[M].push_back(vtotal);
                vofs[r] = vtotal;
                vtotal += M;
            } else {
                vofs[r] = it->second;
            }
        }
    }
    const size_t csz = p->precision == HBX_F64 ? sizeof(double2) : sizeof(float2);
    int dev = 0;
    HBX_HIP(hipGetDevice(&dev));
    int maxlds = 0;
    HBX_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeSharedMemPerBlockOptin, dev));
    if (maxlds <= 0) HBX_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if ((size_t)maxM * csz + (size_t)tw_lds_entries(ilog2(maxM)) * sizeof(double2) > (size_t)maxlds)
        return set_error(HBX_EUNSUPPORTED, "ring FFT of size " + std::to_string(maxM) +
                                               " does not fit in LDS (" + std::to_string(maxlds) +
                                               " bytes) at this precision");
    p->twn = maxM;
    p->twshift = ilog2(maxM);
    p->vtotal = vtotal;
    auto dalloc = [&](void **ptr, size_t bytes) -> int {
        if (bytes == 0) bytes = 16;
        hipError_t e = hipMalloc(ptr, bytes);
        if (e != hipSuccess) return set_error(HBX_ENOMEM, "hipMalloc failed for ring FFT tables");
        p->allocations.push_back(*ptr);
        p->device_bytes += bytes;
        return 0;
    };
    int rc;
    if ((rc = dalloc((void **)&p->d_vofs, nr * sizeof(int64_t)))) return rc;
    HBX_HIP(hipMemcpy(p->d_vofs, vofs.data(), nr * sizeof(int64_t), hipMemcpyHostToDevice));
    // FP64 twiddles for every plan (the LDS tables are FP64 at both precisions)
    if ((rc = dalloc((void **)&p->d_tw64, maxM * sizeof(double2)))) return rc;
    hipLaunchKernelGGL(twiddle_kernel<double>, dim3((maxM + 255) / 256), dim3(256), 0, 0, p->d_tw64, maxM);
    HBX_HIP(hipGetLastError());
    if (p->precision == HBX_F64) {
        if ((rc = dalloc((void **)&p->d_V64, vtotal * sizeof(double2)))) return rc;
        for (auto &kv : hs_by_M)
            if ((rc = build_chirp_spectra<double>(p, kv.second, ofs_by_M[kv.first], kv.first, p->d_V64, p->d_tw64)))
                return rc;
    } else {
        if ((rc = dalloc((void **)&p->d_V32, vtotal * sizeof(float2)))) return rc;
        for (auto &kv : hs_by_M)
            if ((rc = build_chirp_spectra<float>(p, kv.second, ofs_by_M[kv.first], kv.first, p->d_V32, p->d_tw64)))
                return rc;
    }
    // size classes
    for (auto &kv : cls) {
        FftClass c;
        c.M = kv.first;
        c.log2M = ilog2(c.M);
        c.threads = class_threads(c.M);
        c.nb = class_nb(c.M);
        c.nrings = (int)kv.second.size();
        if ((rc = dalloc((void **)&c.d_rings, kv.second.size() * sizeof(int32_t)))) return rc;
        HBX_HIP(hipMemcpy(c.d_rings, kv.second.data(), kv.second.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        c.lds_bytes64 = ring_lds_bytes<double>(c.M, p->twshift);
        c.lds_bytes32 = ring_lds_bytes<float>(c.M, p->twshift);
        p->classes.push_back(c);
    }
    // raise the dynamic-LDS limit of every kernel instance we may launch
    for (const auto &c : p->classes) {
        const size_t lds = p->precision == HBX_F64 ? c.lds_bytes64 : c.lds_bytes32;
        if (lds <= 65536) continue;
        HBX_NB_SWITCH(c.nb, {
            if (p->precision == HBX_F64) {
                if ((rc = allow_lds<double>(ring_analysis_kernel<double, double, NBV>, lds))) return rc;
                if ((rc = allow_lds<double>(ring_analysis_kernel<double, float, NBV>, lds))) return rc;
                if ((rc = allow_lds<double>(ring_synthesis_kernel<double, double, NBV>, lds))) return rc;
                if ((rc = allow_lds<double>(ring_synthesis_kernel<double, float, NBV>, lds))) return rc;
            } else {
                if ((rc = allow_lds<float>(ring_analysis_kernel<float, double, NBV>, lds))) return rc;
                if ((rc = allow_lds<float>(ring_analysis_kernel<float, float, NBV>, lds))) return rc;
                if ((rc = allow_lds<float>(ring_synthesis_kernel<float, double, NBV>, lds))) return rc;
                if ((rc = allow_lds<float>(ring_synthesis_kernel<float, float, NBV>, lds))) return rc;
            }
        });
    }
    HBX_HIP(hipDeviceSynchronize());
    return 0;
}

template <typename T, typename TIn>
static int ring_analysis_t(const Plan *p, int nmaps, const TIn *maps, int64_t map_stride,
                           c2<T> *phase, const double *scale, const double2 *tw, const c2<T> *V,
                           hipStream_t s) {
    const int64_t pstride = phase_elems_per_map(p);
    timer_mark(HBX_TIMER_RING_ANALYSIS, true, s);
    for (const auto &c : p->classes) {
        const size_t lds = ring_lds_bytes<T>(c.M, p->twshift);
        dim3 grid((unsigned)c.nrings, (unsigned)nmaps);
        HBX_NB_SWITCH(c.nb, {
            hipLaunchKernelGGL((ring_analysis_kernel<T, TIn, NBV>), grid, dim3(c.threads), lds, s, maps,
                               map_stride, phase, pstride, c.d_rings, p->d_ring_start, p->d_nphi, p->d_mcut,
                               p->d_phi0, scale, tw, p->twshift, V, p->d_vofs, c.M, c.log2M, p->mmax,
                               p->nrings);
        });
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_error(e, "ring_analysis_kernel launch");
    }
    timer_mark(HBX_TIMER_RING_ANALYSIS, false, s);
    return 0;
}

template <typename T, typename TOut>
static int ring_synthesis_t(const Plan *p, int nmaps, const c2<T> *phase, TOut *maps,
                            int64_t map_stride, const double *scale, const double2 *tw,
                            const c2<T> *V, hipStream_t s) {
    const int64_t pstride = phase_elems_per_map(p);
    timer_mark(HBX_TIMER_RING_SYNTHESIS, true, s);
    for (const auto &c : p->classes) {
        const size_t lds = ring_lds_bytes<T>(c.M, p->twshift);
        dim3 grid((unsigned)c.nrings, (unsigned)nmaps);
        HBX_NB_SWITCH(c.nb, {
            hipLaunchKernelGGL((ring_synthesis_kernel<T, TOut, NBV>), grid, dim3(c.threads), lds, s, phase,
                               pstride, maps, map_stride, c.d_rings, p->d_ring_start, p->d_nphi, p->d_mcut,
                               p->d_phi0, scale, tw, p->twshift, V, p->d_vofs, c.M, c.log2M, p->mmax,
                               p->nrings);
        });
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_error(e, "ring_synthesis_kernel launch");
    }
    timer_mark(HBX_TIMER_RING_SYNTHESIS, false, s);
    return 0;
}

int launch_ring_analysis(const Plan *p, int nmaps, const void *maps, int map_dtype,
                         int64_t map_stride, void *phase, const double *scale, hipStream_t s) {
    if (p->precision == HBX_F64) {
        auto *ph = static_cast<double2 *>(phase);
        if (map_dtype == HBX_DT_F64)
            return ring_analysis_t<double, double>(p, nmaps, (const double *)maps, map_stride, ph, scale, p->d_tw64, p->d_V64, s);
        return ring_analysis_t<double, float>(p, nmaps, (const float *)maps, map_stride, ph, scale, p->d_tw64, p->d_V64, s);
    }
    auto *ph = static_cast<float2 *>(phase);
    if (map_dtype == HBX_DT_F64)
        return ring_analysis_t<float, double>(p, nmaps, (const double *)maps, map_stride, ph, scale, p->d_tw64, p->d_V32, s);
    return ring_analysis_t<float, float>(p, nmaps, (const float *)maps, map_stride, ph, scale, p->d_tw64, p->d_V32, s);
}

int launch_ring_synthesis(const Plan *p, int nmaps, const void *phase, void *maps, int map_dtype,
                          int64_t map_stride, const double *scale, hipStream_t s) {
    if (p->precision == HBX_F64) {
        auto *ph = static_cast<const double2 *>(phase);
        if (map_dtype == HBX_DT_F64)
            return ring_synthesis_t<double, double>(p, nmaps, ph, (double *)maps, map_stride, scale, p->d_tw64, p->d_V64, s);
        return ring_synthesis_t<double, float>(p, nmaps, ph, (float *)maps, map_stride, scale, p->d_tw64, p->d_V64, s);
    }
    auto *ph = static_cast<const float2 *>(phase);
    if (map_dtype == HBX_DT_F64)
        return ring_synthesis_t<float, double>(p, nmaps, ph, (double *)maps, map_stride, scale, p->d_tw64, p->d_V32, s);
    return ring_synthesis_t<float, float>(p, nmaps, ph, (float *)maps, map_stride, scale, p->d_tw64, p->d_V32, s);
}

}  // namespace hbx
