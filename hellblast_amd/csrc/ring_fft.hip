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
// Twiddles in LDS, indexed relative to the class FFT size M:
//   RingTw<T>::get(j) = e^{-2 pi i j / M},  0 <= j < M.
// FP32: the first octant (j <= M/8), correctly rounded (strided from the
// plan's FP32 table of e^{-2 pi i J / 2^twshift}, rounded once from FP64);
// the rest of the circle follows by exact swaps and negations, so every
// twiddle an FP32 transform uses is correctly rounded (the FP32 ring stage's
// error budget is set by them).  M/8 + 1 entries: 8 KB for 8192-point rings.
// FP64: a two-level table over 2^twshift, w^J = coarse[J >> f] * fine[J &
// (2^f - 1)], f = min(7, twshift), the product formed in FP64.
// ---------------------------------------------------------------------------
constexpr int TW_FINE_LOG = 7;
inline int tw_fine_log(int twshift) { return twshift < TW_FINE_LOG ? twshift : TW_FINE_LOG; }
__device__ __forceinline__ int tw_fine_log_d(int twshift) { return min(twshift, TW_FINE_LOG); }
inline int tw_lds_entries(int twshift) {
    const int f = tw_fine_log(twshift);
    return (1 << (twshift - f)) + (1 << f);
}

template <typename T> struct RingTw;

template <> struct RingTw<float> {
    const float2 *tab;
    int qlog;   // log2(M / 4); negative for M <= 2 (only j = 0 occurs)
    static int entries(int M) { return M >= 8 ? M / 8 + 1 : 1; }
    static size_t lds_bytes(int M, int) { return (size_t)entries(M) * sizeof(float2); }
    __device__ static void stage(void *lds, const void *gtab, int M, int log2M, int twshift) {
        float2 *t = static_cast<float2 *>(lds);
        const float2 *g = static_cast<const float2 *>(gtab);
        const int n = M >= 8 ? M / 8 + 1 : 1, sh = twshift - log2M;
        for (int i = threadIdx.x; i < n; i += blockDim.x) t[i] = g[i << sh];
    }
    __device__ RingTw(const void *lds, int log2M, int) : tab(static_cast<const float2 *>(lds)), qlog(log2M - 2) {}
    __device__ __forceinline__ float2 get(int j) const {
        if (qlog < 0) return make_float2(1.f, 0.f);
        const int q = j >> qlog, jr = j & ((1 << qlog) - 1);        // j = q M/4 + jr
        const bool sw = qlog > 0 && jr > (1 << (qlog - 1));           // second octant
        const float2 t = tab[sw ? (1 << qlog) - jr : jr];
        float2 w = sw ? make_float2(-t.y, -t.x) : t;                  // e^{-i pi/2} conj(t)
        if (q & 1) w = make_float2(w.y, -w.x);                        // times -i
        if (q & 2) w = make_float2(-w.x, -w.y);                       // times -1
        return w;
    }
};

template <> struct RingTw<double> {
    const double2 *tab;
    int f, nc, sh;
    static size_t lds_bytes(int, int twshift) { return (size_t)tw_lds_entries(twshift) * sizeof(double2); }
    __device__ static void stage(void *lds, const void *gtab, int, int, int twshift) {
        double2 *t = static_cast<double2 *>(lds);
        const double2 *g = static_cast<const double2 *>(gtab);
        const int f = tw_fine_log_d(twshift), nc = 1 << (twshift - f), nf = 1 << f;
        for (int i = threadIdx.x; i < nc + nf; i += blockDim.x) t[i] = (i < nc) ? g[i << f] : g[i - nc];
    }
    __device__ RingTw(const void *lds, int log2M, int twshift)
        : tab(static_cast<const double2 *>(lds)), f(tw_fine_log_d(twshift)), nc(1 << (twshift - f)),
          sh(twshift - log2M) {}
    __device__ __forceinline__ double2 get(int j) const {
        const int J = j << sh;
        const double2 a = tab[J >> f], b = tab[nc + (J & ((1 << f) - 1))];
        return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
    }
};

// LDS bytes of a ring workgroup: M data elements + the twiddle table
template <typename T> inline size_t ring_lds_bytes(int M, int twshift) {
    return (size_t)M * sizeof(c2<T>) + RingTw<T>::lds_bytes(M, twshift);
}

// ---------------------------------------------------------------------------
// in-LDS FFT of size M (power of two), forward (e^{-i}) or inverse (e^{+i},
// unnormalised).  Stockham radix-4 passes plus one radix-2 pass when log2 M
// is odd; each pass stages its inputs in registers so one buffer suffices.
// tw: the staged LDS twiddle table (visible to every thread).
// ---------------------------------------------------------------------------
template <typename T, int NB>
__device__ __forceinline__ void lds_fft(c2<T> *buf, int M, int log2M, const RingTw<T> &tw, bool inv) {
    const int nt = blockDim.x, tid = threadIdx.x;
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
        const int sh = log2M - lp - 2;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int i = tid + b * nt;
            if (i < Q) {
                const int k = i & (p - 1);
                c2<T> w1 = tw.get(k << sh), w2 = tw.get((2 * k) << sh), w3 = tw.get((3 * k) << sh);
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
#pragma unroll
        for (int b = 0; b < 2 * NB; ++b) {
            const int i = tid + b * nt;
            if (i < H) {
                c2<T> w = tw.get(i);
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

// chirp c_j = exp(i pi j^2 / h), with j^2 reduced mod 2h exactly (plan time)
template <typename T>
__device__ __forceinline__ c2<T> chirp(int j, int h) {
    const long long q = ((long long)j * j) % (2LL * h);
    return expipi<T>((T)q / (T)h);
}

template <typename T> __device__ __forceinline__ T load_as(const double *p) { return (T)(*p); }
template <typename T> __device__ __forceinline__ T load_as(const float *p) { return (T)(*p); }

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
template <typename T> __device__ __forceinline__ c2<T> round_c(double2 a) { return mk<T>((T)a.x, (T)a.y); }

// Per-ring rotations (plan time, FP64), nt = the ring's workgroup size:
//   rot[0][r] = e^{i nt phi0}   rot[1][r] = e^{i n phi0}   rot[2][r] = e^{i h phi0}
//   rot[3][r] = e^{-2 pi i nt / n}
// The ring kernels advance e^{-+ i m phi0} and e^{-+ 2 pi i m / n} from one
// sincos per thread by these FP64 steps (at most a few dozen per thread).
constexpr int NROT = 4;

// ---------------------------------------------------------------------------
// analysis: maps -> phase
// V + vofs[r]: chirp spectrum (M entries) then the chirp c_j, j < h
// ---------------------------------------------------------------------------
template <typename T, typename TIn, int NB>
__global__ void __launch_bounds__(1024)
ring_analysis_kernel(const TIn *__restrict__ maps, int64_t map_stride, c2<T> *__restrict__ phase,
                     int64_t phase_stride, const int32_t *__restrict__ rings,
                     const int64_t *__restrict__ ring_start, const int32_t *__restrict__ nphi,
                     const int32_t *__restrict__ mcut, const double *__restrict__ phi0,
                     const double *__restrict__ scale, const double2 *__restrict__ rot, const void *tw,
                     int twshift, const c2<T> *__restrict__ V, const int64_t *__restrict__ vofs, int M, int log2M,
                     int mmax, int nr) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    c2<T> *buf = reinterpret_cast<c2<T> *>(smem_raw);
    const int r = rings[blockIdx.x];
    const int b = blockIdx.y;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int n = nphi[r], h = n >> 1;
    const bool pow2 = (h == M);
    const TIn *f = maps + (int64_t)b * map_stride + ring_start[r];
    RingTw<T>::stage(buf + M, tw, M, log2M, twshift);
    const RingTw<T> twl(buf + M, log2M, twshift);
    const c2<T> *Vr = pow2 ? nullptr : V + vofs[r];
    const c2<T> *Cr = pow2 ? nullptr : Vr + M;

    for (int j = tid; j < M; j += nt) {
        c2<T> z = mk<T>(0, 0);
        if (j < h) {
            z = mk<T>(load_as<T>(f + 2 * j), load_as<T>(f + 2 * j + 1));
            if (!pow2) z = cmul<T>(z, cconj<T>(Cr[j]));
        }
        buf[j] = z;
    }
    __syncthreads();
    lds_fft<T, NB>(buf, M, log2M, twl, false);
    if (!pow2) {
        for (int k = tid; k < M; k += nt) buf[k] = cmul<T>(buf[k], Vr[k]);
        __syncthreads();
        lds_fft<T, NB>(buf, M, log2M, twl, true);
        const T invM = (T)1 / (T)M;
        for (int k = tid; k < h; k += nt) buf[k] = cscale<T>(cmul<T>(buf[k], cconj<T>(Cr[k])), invM);
        __syncthreads();
    }
    // untangle: X[m] = E + W_m O with E = (Z[m] + conj Z[-m]) / 2, O = (Z[m] -
    // conj Z[-m]) / 2i (Z of period h), W_m = e^{-2 pi i m / n}; then
    // F(m) = scale e^{-i m phi0} X[m] for m < mcut.
    const T sc = (T)(scale ? scale[r] : 1.0);
    const int cut = mcut[r];
    const double ph0 = phi0 ? phi0[r] : 0.0;
    const bool rotp = (ph0 != 0.0);
    double2 W, P = make_double2(1.0, 0.0);
    {
        double s, c;
        sincospi(-2.0 * (double)(tid % n) / (double)n, &s, &c);
        W = make_double2(c, s);
        if (rotp) {
            sincos(-(double)tid * ph0, &s, &c);
            P = make_double2(c, s);
        }
    }
    const double2 Wst = rot[3 * nr + r], Pst = zconj(rot[r]);
    const int hstep = nt % h;
    int ia = tid % h;
    c2<T> *out = phase + (int64_t)b * phase_stride;
    for (int m = tid; m <= mmax; m += nt) {
        c2<T> F = mk<T>(0, 0);
        if (m < cut) {
            const int ib = (ia == 0) ? 0 : h - ia;
            const c2<T> za = buf[ia], zb = cconj<T>(buf[ib]);
            const c2<T> A = cscale<T>(cadd<T>(za, zb), (T)0.5);
            const c2<T> D = cscale<T>(csub<T>(za, zb), (T)0.5);
            const c2<T> X = cadd<T>(A, cmul<T>(round_c<T>(W), cmul_mi<T>(D)));
            F = cscale<T>(X, sc);
            if (rotp) F = cmul<T>(F, round_c<T>(P));
        }
        out[((int64_t)(m >> 3) * nr + r) * MB + (m & 7)] = F;
        W = zmul(W, Wst);
        if (rotp) P = zmul(P, Pst);
        ia += hstep;
        if (ia >= h) ia -= h;
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
                      const double *__restrict__ scale, const double2 *__restrict__ rot, const void *tw,
                      int twshift, const c2<T> *__restrict__ V, const int64_t *__restrict__ vofs, int M,
                      int log2M, int mmax, int nr) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    c2<T> *buf = reinterpret_cast<c2<T> *>(smem_raw);
    const int r = rings[blockIdx.x];
    const int b = blockIdx.y;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int n = nphi[r], h = n >> 1;
    const bool pow2 = (h == M);
    const int top = min(mmax, mcut[r] - 1);
    RingTw<T>::stage(buf + M, tw, M, log2M, twshift);
    const RingTw<T> twl(buf + M, log2M, twshift);
    const c2<T> *Vr = pow2 ? nullptr : V + vofs[r];
    const c2<T> *Cr = pow2 ? nullptr : Vr + M;
    const double ph0 = phi0 ? phi0[r] : 0.0;
    const bool rotp = (ph0 != 0.0);
    const c2<T> *G = phase + (int64_t)b * phase_stride;
    const double2 Rn = rot[nr + r];

    // X[k] = sum_{m = k mod n} G'(m) + sum_{m = -k mod n, m >= 1} conj G'(m),
    // G'(m) = G(m) e^{i m phi0} (imaginary part of G(0) dropped); Rk = e^{i k phi0}
    auto xsum = [&](int k, double2 Rk) -> c2<T> {
        c2<T> acc = mk<T>(0, 0);
        double2 q = Rk;
        for (int m = k; m <= top; m += n) {
            c2<T> g = G[((int64_t)(m >> 3) * nr + r) * MB + (m & 7)];
            if (m == 0) g.y = 0;
            if (rotp) {
                g = cmul<T>(g, round_c<T>(q));
                q = zmul(q, Rn);
            }
            acc = cadd<T>(acc, g);
        }
        q = zmul(zconj(Rk), Rn);
        for (int m = (k == 0 ? n : n - k); m <= top; m += n) {
            c2<T> g = G[((int64_t)(m >> 3) * nr + r) * MB + (m & 7)];
            if (rotp) {
                g = cmul<T>(g, round_c<T>(q));
                q = zmul(q, Rn);
            }
            acc = cadd<T>(acc, cconj<T>(g));
        }
        return acc;
    };
    // per thread, k = tid + t nt: R = e^{i k phi0}, W = e^{2 pi i k / n}
    double2 R = make_double2(1.0, 0.0), W;
    {
        double s, c;
        sincospi(2.0 * (double)(tid % n) / (double)n, &s, &c);
        W = make_double2(c, s);
        if (rotp) {
            sincos((double)tid * ph0, &s, &c);
            R = make_double2(c, s);
        }
    }
    const double2 S = rot[r], Rh = rot[2 * nr + r], Wst = zconj(rot[3 * nr + r]);
    for (int k = tid; k <= (h >> 1); k += nt) {
        const int k2 = h - k;
        const c2<T> Xa = xsum(k, R), Xb = xsum(k2, rotp ? zmul(Rh, zconj(R)) : R);
        const c2<T> Wk = round_c<T>(W);
        const c2<T> Wk2 = mk<T>(-Wk.x, Wk.y);        // e^{2 pi i (h - k)/n} = -conj W_k
        {
            const c2<T> s = cadd<T>(Xa, cconj<T>(Xb)), d = csub<T>(Xa, cconj<T>(Xb));
            c2<T> z = cadd<T>(s, cmul_pi<T>(cmul<T>(Wk, d)));
            if (!pow2) z = cconj<T>(cmul<T>(z, Cr[k]));
            buf[k] = z;
        }
        if (k != 0 && k2 != k) {
            const c2<T> s = cadd<T>(Xb, cconj<T>(Xa)), d = csub<T>(Xb, cconj<T>(Xa));
            c2<T> z = cadd<T>(s, cmul_pi<T>(cmul<T>(Wk2, d)));
            if (!pow2) z = cconj<T>(cmul<T>(z, Cr[k2]));
            buf[k2] = z;
        }
        W = zmul(W, Wst);
        if (rotp) R = zmul(R, S);
    }
    if (!pow2)
        for (int j = h + tid; j < M; j += nt) buf[j] = mk<T>(0, 0);
    __syncthreads();
    if (pow2) {
        lds_fft<T, NB>(buf, M, log2M, twl, true);
    } else {
        lds_fft<T, NB>(buf, M, log2M, twl, false);
        for (int k = tid; k < M; k += nt) buf[k] = cmul<T>(buf[k], Vr[k]);
        __syncthreads();
        lds_fft<T, NB>(buf, M, log2M, twl, true);
        const T invM = (T)1 / (T)M;
        for (int j = tid; j < h; j += nt) buf[j] = cscale<T>(cmul<T>(Cr[j], cconj<T>(buf[j])), invM);
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
// (same conventions as ring_analysis_kernel), on the FP64 matrix cores: per
// ring and group of 16 maps, v_mfma_f64_16x16x4f64 contracts 4 pixels per
// instruction,
//   D[row = 2 m + c][col = map] += A[row][k] B[k][col],
//   A[2 m][k] = cos(2 pi m j_k / n), A[2 m + 1][k] = -sin(2 pi m j_k / n),
//   B[k][col] = f_col(j_k).
// Each of 4 waves takes a quarter of the ring, staging 64 pixels of the 16
// maps at a time in LDS (coalesced loads); its lanes carry e^{-2 pi i m j/n}
// for their (row, pixel) by an FP64 recurrence over 4-pixel steps, anchored
// by one sincospi per 64 pixels.  Lane layout of the 16x16x4 f64
// instruction: A lane = row + 16 k, B lane = col + 16 k, D lane = col + 16
// (row % 4), reg = row / 4.
// ---------------------------------------------------------------------------
typedef double dx4 __attribute__((ext_vector_type(4)));
constexpr int RB_WAVES = 4;
constexpr int RB_CHUNK = 64;

template <typename TIn>
__global__ void __launch_bounds__(RB_WAVES * 64)
ring_band_kernel(const TIn *__restrict__ maps, int64_t map_stride, double2 *__restrict__ band, int64_t band_stride,
                 const int64_t *__restrict__ ring_start, const int32_t *__restrict__ nphi,
                 const int32_t *__restrict__ mcut, const double *__restrict__ phi0, const double *__restrict__ scale,
                 int mtop, int nr, int nmaps) {
    __shared__ TIn stg[RB_WAVES][16][RB_CHUNK + 1];
    __shared__ double dsum[RB_WAVES][4][64];
    const int r = blockIdx.x, g = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int n = nphi[r];
    const TIn *f = maps + ring_start[r];
    const int per = ((n + RB_WAVES - 1) / RB_WAVES + 3) / 4 * 4;   // pixels per wave, a multiple of 4
    const int jb = w * per, je = min(n, jb + per);
    const int row = lane & 15, k = lane >> 4;
    const int m = row >> 1, im = row & 1;
    double sr, si;                                  // e^{-2 pi i 4 m / n}: the 4-pixel step
    sincospi(-2.0 * (double)((4 * m) % n) / (double)n, &si, &sr);
    dx4 D = {0.0, 0.0, 0.0, 0.0};
    // the 16 maps' pixels of the next chunk are loaded while this one computes
    TIn nx[16];
    auto load = [&](int c0) {
        const int j = c0 + lane;
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            const int bb = g * 16 + b;
            nx[b] = (j < je && bb < nmaps) ? f[(int64_t)bb * map_stride + j] : (TIn)0;
        }
    };
    if (jb < je) load(jb);
    for (int c0 = jb; c0 < je; c0 += RB_CHUNK) {
#pragma unroll
        for (int b = 0; b < 16; ++b) stg[w][b][lane] = nx[b];
        if (c0 + RB_CHUNK < je) load(c0 + RB_CHUNK);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double cr, ci;                              // e^{-2 pi i m j / n}, j = c0 + k
        sincospi(-2.0 * (double)((m * (c0 + k)) % n) / (double)n, &ci, &cr);
#pragma unroll 4
        for (int s = 0; s < RB_CHUNK / 4; ++s) {
            D = __builtin_amdgcn_mfma_f64_16x16x4f64(im ? ci : cr, (double)stg[w][row][4 * s + k], D, 0, 0, 0);
            const double t = cr * sr - ci * si;
            ci = cr * si + ci * sr;
            cr = t;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) dsum[w][q][lane] = D[q];
    __syncthreads();
    if (tid < 16 * M_LO) {
        const int mm = tid >> 4, b = tid & 15;       // output (m, map)
        const int bb = g * 16 + b;
        if (mm <= mtop && bb < nmaps) {
            // D[row][col]: lane = col + 16 (row % 4), reg = row / 4
            double x = 0.0, y = 0.0;
            const int r0 = 2 * mm, r1 = 2 * mm + 1;
#pragma unroll
            for (int ww = 0; ww < RB_WAVES; ++ww) {
                x += dsum[ww][r0 >> 2][b + 16 * (r0 & 3)];
                y += dsum[ww][r1 >> 2][b + 16 * (r1 & 3)];
            }
            double2 F = make_double2(0.0, 0.0);
            if (mm < mcut[r]) {
                const double sc = scale ? scale[r] : 1.0;
                x *= sc;
                y *= sc;
                const double ph = phi0 ? phi0[r] : 0.0;
                if (ph != 0.0) {
                    double sn, cs;
                    sincos(-(double)mm * ph, &sn, &cs);
                    F = make_double2(x * cs - y * sn, x * sn + y * cs);
                } else {
                    F = make_double2(x, y);
                }
            }
            band[(int64_t)bb * band_stride + ((int64_t)(mm >> 3) * nr + r) * MB + (mm & 7)] = F;
        }
    }
}

int launch_ring_band(const Plan *p, int nmaps, const void *maps, int map_dtype, int64_t map_stride, void *band,
                     const double *scale, hipStream_t s) {
    const int mtop = std::min(M_LO - 1, p->mmax);
    dim3 grid(p->nrings, (nmaps + 15) / 16);
    if (map_dtype == HBX_DT_F64)
        hipLaunchKernelGGL(ring_band_kernel<double>, grid, dim3(RB_WAVES * 64), 0, s, (const double *)maps,
                           map_stride, (double2 *)band, band_phase_elems(p), p->d_ring_start, p->d_nphi, p->d_mcut,
                           p->d_phi0, scale, mtop, p->nrings, nmaps);
    else
        hipLaunchKernelGGL(ring_band_kernel<float>, grid, dim3(RB_WAVES * 64), 0, s, (const float *)maps,
                           map_stride, (double2 *)band, band_phase_elems(p), p->d_ring_start, p->d_nphi, p->d_mcut,
                           p->d_phi0, scale, mtop, p->nrings, nmaps);
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

// V_h = FFT_M(v), v_j = c_j (0 <= j < h), v_{M-j} = c_j (1 <= j < h), else 0,
// stored at V + ofs followed by the chirp c_j itself (j < h)
template <typename T, int NB>
__global__ void __launch_bounds__(1024)
chirp_spectrum_kernel(const int32_t *__restrict__ hs, const int64_t *__restrict__ ofs, c2<T> *V, const void *tw,
                      int twshift, int M, int log2M) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    c2<T> *buf = reinterpret_cast<c2<T> *>(smem_raw);
    const int h = hs[blockIdx.x];
    RingTw<T>::stage(buf + M, tw, M, log2M, twshift);
    const RingTw<T> twl(buf + M, log2M, twshift);
    c2<T> *out = V + ofs[blockIdx.x];
    for (int j = threadIdx.x; j < M; j += blockDim.x) {
        c2<T> v = mk<T>(0, 0);
        if (j < h) {
            v = chirp<T>(j, h);
            out[M + j] = v;
        } else if (j > M - h) {
            v = chirp<T>(M - j, h);
        }
        buf[j] = v;
    }
    __syncthreads();
    lds_fft<T, NB>(buf, M, log2M, twl, false);
    for (int j = threadIdx.x; j < M; j += blockDim.x) out[j] = buf[j];
}

__global__ void round_to_f32_kernel(const double2 *__restrict__ src, float2 *__restrict__ dst, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = make_float2((float)src[i].x, (float)src[i].y);
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

// chirp spectra and chirps of the rings of one size class M, computed at
// precision T into V (offsets ofs_host, one entry per distinct h)
template <typename T>
static int build_chirp_spectra(Plan *p, const std::vector<int32_t> &hs_host, const std::vector<int64_t> &ofs_host,
                               int M, c2<T> *V, const void *tw) {
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

int plan_build_fft_tables(Plan *p, const std::vector<int32_t> &nphi, const double *phi0) {
    const int nr = p->nrings;
    // class per ring
    std::map<int, std::vector<int32_t>> cls;            // M -> rings
    std::map<int, int64_t> vofs_of_h;                   // h -> offset (non-pow2)
    std::map<int, std::vector<int32_t>> hs_by_M;        // M -> distinct h
    std::map<int, std::vector<int64_t>> ofs_by_M;
    std::vector<int64_t> vofs(nr, -1);
    std::vector<int> ringM(nr);
    int maxM = 4;
    int64_t vtotal = 0;
    for (int r = 0; r < nr; ++r) {
        const int h = nphi[r] / 2;
        const bool pow2 = (h & (h - 1)) == 0;
        const int M = pow2 ? h : (1 << ilog2(2 * (int64_t)h - 1));
        ringM[r] = M;
        cls[M].push_back(r);
        maxM = std::max(maxM, M);
        if (!pow2) {
            auto it = vofs_of_h.find(h);
            if (it == vofs_of_h.end()) {
                vofs_of_h[h] = vtotal;
                hs_by_M[M].push_back(h);
                ofs_by_M[M].push_back(vtotal);
                vofs[r] = vtotal;
                vtotal += M + h;                       // spectrum, then chirp
            } else {
                vofs[r] = it->second;
            }
        }
    }
    int dev = 0;
    HBX_HIP(hipGetDevice(&dev));
    int maxlds = 0;
    HBX_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeSharedMemPerBlockOptin, dev));
    if (maxlds <= 0) HBX_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    const int twshift = ilog2(maxM);
    const size_t need = p->precision == HBX_F64 ? ring_lds_bytes<double>(maxM, twshift)
                                                : ring_lds_bytes<float>(maxM, twshift);
    if (need > (size_t)maxlds)
        return set_error(HBX_EUNSUPPORTED, "ring FFT of size " + std::to_string(maxM) +
                                               " does not fit in LDS (" + std::to_string(maxlds) +
                                               " bytes) at this precision");
    p->twn = maxM;
    p->twshift = twshift;
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
    // per-ring rotation steps (long double on the host, rounded to FP64)
    {
        std::vector<double2> rot((size_t)NROT * nr);
        const long double pi = 3.141592653589793238462643383279502884L;
        for (int r = 0; r < nr; ++r) {
            const int n = nphi[r], h = n / 2, nt = class_threads(ringM[r]);
            const long double ph = phi0 ? (long double)phi0[r] : 0.0L;
            auto e = [](long double a) { return make_double2((double)cosl(a), (double)sinl(a)); };
            rot[r] = e(nt * ph);
            rot[(size_t)nr + r] = e(n * ph);
            rot[2 * (size_t)nr + r] = e(h * ph);
            rot[3 * (size_t)nr + r] = e(-2.0L * pi * (long double)(nt % n) / (long double)n);
        }
        if ((rc = dalloc((void **)&p->d_rot, rot.size() * sizeof(double2)))) return rc;
        HBX_HIP(hipMemcpy(p->d_rot, rot.data(), rot.size() * sizeof(double2), hipMemcpyHostToDevice));
    }
    // FP64 twiddles for every plan; FP32 plans also get them rounded once
    if ((rc = dalloc((void **)&p->d_tw64, maxM * sizeof(double2)))) return rc;
    hipLaunchKernelGGL(twiddle_kernel<double>, dim3((maxM + 255) / 256), dim3(256), 0, 0, p->d_tw64, maxM);
    HBX_HIP(hipGetLastError());
    if (p->precision == HBX_F32) {
        if ((rc = dalloc((void **)&p->d_tw32, maxM * sizeof(float2)))) return rc;
        hipLaunchKernelGGL(twiddle_kernel<float>, dim3((maxM + 255) / 256), dim3(256), 0, 0, p->d_tw32, maxM);
        HBX_HIP(hipGetLastError());
    }
    if (p->precision == HBX_F64) {
        if ((rc = dalloc((void **)&p->d_V64, vtotal * sizeof(double2)))) return rc;
        for (auto &kv : hs_by_M)
            if ((rc = build_chirp_spectra<double>(p, kv.second, ofs_by_M[kv.first], kv.first, p->d_V64, p->d_tw64)))
                return rc;
    } else {
        if ((rc = dalloc((void **)&p->d_V32, vtotal * sizeof(float2)))) return rc;
        // spectra and chirps in FP64, rounded once, where the FP64 FFT fits in LDS
        double2 *tmp = nullptr;
        if (vtotal > 0 && hipMalloc(&tmp, vtotal * sizeof(double2)) != hipSuccess)
            return set_error(HBX_ENOMEM, "hipMalloc failed for the FP64 chirp spectra");
        for (auto &kv : hs_by_M) {
            if (ring_lds_bytes<double>(kv.first, twshift) <= (size_t)maxlds)
                rc = build_chirp_spectra<double>(p, kv.second, ofs_by_M[kv.first], kv.first, tmp, p->d_tw64);
            else
                rc = build_chirp_spectra<float>(p, kv.second, ofs_by_M[kv.first], kv.first, p->d_V32, p->d_tw32);
            if (rc) break;
            // round this class's entries (every distinct h of the class)
            for (size_t i = 0; i < kv.second.size() && ring_lds_bytes<double>(kv.first, twshift) <= (size_t)maxlds;
                 ++i) {
                const int64_t o = ofs_by_M[kv.first][i], len = kv.first + kv.second[i];
                hipLaunchKernelGGL(round_to_f32_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, 0,
                                   tmp + o, p->d_V32 + o, len);
            }
        }
        hipError_t e = hipDeviceSynchronize();
        if (tmp) (void)hipFree(tmp);
        if (rc) return rc;
        if (e != hipSuccess) return hip_error(e, "FP32 chirp spectra");
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
                           c2<T> *phase, const double *scale, const void *tw, const c2<T> *V,
                           hipStream_t s) {
    const int64_t pstride = phase_elems_per_map(p);
    timer_mark(HBX_TIMER_RING_ANALYSIS, true, s);
    for (const auto &c : p->classes) {
        const size_t lds = ring_lds_bytes<T>(c.M, p->twshift);
        dim3 grid((unsigned)c.nrings, (unsigned)nmaps);
        HBX_NB_SWITCH(c.nb, {
            hipLaunchKernelGGL((ring_analysis_kernel<T, TIn, NBV>), grid, dim3(c.threads), lds, s, maps,
                               map_stride, phase, pstride, c.d_rings, p->d_ring_start, p->d_nphi, p->d_mcut,
                               p->d_phi0, scale, p->d_rot, tw, p->twshift, V, p->d_vofs, c.M, c.log2M,
                               p->mmax, p->nrings);
        });
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_error(e, "ring_analysis_kernel launch");
    }
    timer_mark(HBX_TIMER_RING_ANALYSIS, false, s);
    return 0;
}

template <typename T, typename TOut>
static int ring_synthesis_t(const Plan *p, int nmaps, const c2<T> *phase, TOut *maps,
                            int64_t map_stride, const double *scale, const void *tw,
                            const c2<T> *V, hipStream_t s) {
    const int64_t pstride = phase_elems_per_map(p);
    timer_mark(HBX_TIMER_RING_SYNTHESIS, true, s);
    for (const auto &c : p->classes) {
        const size_t lds = ring_lds_bytes<T>(c.M, p->twshift);
        dim3 grid((unsigned)c.nrings, (unsigned)nmaps);
        HBX_NB_SWITCH(c.nb, {
            hipLaunchKernelGGL((ring_synthesis_kernel<T, TOut, NBV>), grid, dim3(c.threads), lds, s, phase,
                               pstride, maps, map_stride, c.d_rings, p->d_ring_start, p->d_nphi, p->d_mcut,
                               p->d_phi0, scale, p->d_rot, tw, p->twshift, V, p->d_vofs, c.M, c.log2M,
                               p->mmax, p->nrings);
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
        return ring_analysis_t<float, double>(p, nmaps, (const double *)maps, map_stride, ph, scale, p->d_tw32, p->d_V32, s);
    return ring_analysis_t<float, float>(p, nmaps, (const float *)maps, map_stride, ph, scale, p->d_tw32, p->d_V32, s);
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
        return ring_synthesis_t<float, double>(p, nmaps, ph, (double *)maps, map_stride, scale, p->d_tw32, p->d_V32, s);
    return ring_synthesis_t<float, float>(p, nmaps, ph, (float *)maps, map_stride, scale, p->d_tw32, p->d_V32, s);
}

}  // namespace hbx
