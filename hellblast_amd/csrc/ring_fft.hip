// Ring phase stage: batched real DFTs over HEALPix rings of uneven length.
//
// Reference: per-ring cupy.fft.fft(ring)/nphi (sht_raw.py:38-43,
// sht_fallback.py:51-56) and cupy.fft.ifft(F)*nphi (isht.py:24), with the
// alias rule of sht_raw.py:51 (ring contributes to m < nphi only).
//
// One workgroup transforms one ring of one map entirely in LDS:
//   * a real ring of length n = 2h is packed as h complex values
//     z_j = x_{2j} + i x_{2j+1} (one half-length complex DFT per ring);
//   * h a power of two  -> one FFT of size M = h;
//   * otherwise         -> Bluestein chirp-z with FFTs of a power-of-two size
//     M >= 2h-1 and a per-length chirp spectrum precomputed at plan time;
//   * the half-length spectrum is untangled into X[k], k <= h, and written as
//     F_r(m) = scale_r * e^{-i m phi0_r} X[m mod n]  for m < mcut_r, else 0,
//     into the m-blocked ring-major phase buffer [map][m/8][ring][m%8].
// Rings are grouped into size classes (power-of-two M, and kind: power of
// two or Bluestein); one launch runs every class of one workgroup size and
// kind, so a ring stage is a handful of launches, each reserving the LDS of
// its largest class.
//
// The FFT engine keeps 16 points per thread in registers: radix-16 passes
// (4 radix-2 levels each, one LDS exchange per pass) and one radix-2/4/8
// pass for the remaining levels, Stockham order, twiddles read from a
// per-pass table of correctly rounded values.
#include <algorithm>
#include <cmath>
#include <map>

#include "hbx_device.h"
#include "hbx_internal.h"

namespace hbx {

// ---------------------------------------------------------------------------
// FFT engine.  Size M = 2^L, M/P threads (one for M < P) hold P points each
// (P = 2^q: 16 in FP32, 8 in FP64).  Pass s has radix r_s (P for the first
// floor(L/q) passes, then 2^(L mod q) if nonzero) and stride
// p_s = r_0 ... r_{s-1}; butterfly i
// (0 <= i < M/r_s), k = i mod p_s, reads x_j = buf[i + j M/r_s], multiplies
// by w_{r_s p_s}^{+-jk}, takes the r_s-point DFT y and writes y_j to
// buf[(i - k) r_s + k + j p_s]  (Stockham autosort: natural order in and
// out).  The first pass (p = 1) needs no twiddles and reads through a
// caller functor; the last writes through one.
//
// LDS index x is padded to x + x/P so that the stride-P writes of the
// first passes spread over the banks: (M + M/P) elements per workgroup.
//
// Twiddle table of a class: for each pass s >= 1, p_s rows of r_s entries,
// row k = w^{0 .. r_s - 1 times k} with w = e^{-2 pi i / (r_s p_s)}, rows
// 16-byte aligned so a thread loads its row with vector loads.  Entries are
// computed in long double on the host and rounded once.
// ---------------------------------------------------------------------------
// log2 of the points per thread: 16 in FP32, 8 in FP64 (a radix-8 pass of
// double-complex values stays near 100 VGPRs, so 4 waves share a SIMD)
template <typename T> constexpr int fft_lp() { return sizeof(T) == 8 ? 3 : 4; }
template <typename T> constexpr int fft_pts() { return 1 << fft_lp<T>(); }

template <typename T> __device__ __forceinline__ int pad_idx(int x) { return x + (x >> fft_lp<T>()); }
template <typename T> inline int64_t pad_len(int64_t M) { return M + (M >> fft_lp<T>()); }

// threads of a ring workgroup of class M (also the ring kernels' stride)
template <typename T> inline int class_threads(int M) { return std::max(64, M / fft_pts<T>()); }

// (one element more than the FFT: the analysis untangle writes X[h] there)
template <typename T> inline size_t ring_lds_bytes(int M) { return (size_t)(pad_len<T>(M) + 1) * sizeof(c2<T>); }

// a * (-i) forward, a * (+i) inverse
template <typename T, bool INV> __device__ __forceinline__ c2<T> mul_wi(c2<T> a) {
    return INV ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
}
// a * (cr + i ci) forward, a * conj(cr + i ci) inverse
template <typename T, bool INV> __device__ __forceinline__ c2<T> mul_c(c2<T> a, T cr, T ci) {
    if (INV) ci = -ci;
    return mk<T>(a.x * cr - a.y * ci, a.x * ci + a.y * cr);
}
// the same by an internal constant of the radix-8/16 DFTs (given in FP64).
// FP32: the constant is carried as hi + lo and both parts are applied, so
// its rounding (the same in every butterfly of every ring) adds no error
// that is coherent across rings.  FP64: plain product.
template <typename T, bool INV> __device__ __forceinline__ c2<T> mul_k(c2<T> a, double crd, double cid) {
    if constexpr (sizeof(T) == 8) {
        return mul_c<T, INV>(a, crd, cid);
    } else {
        const float cr = (float)crd, ci0 = (float)cid;
        const float crl = (float)(crd - (double)cr), cil0 = (float)(cid - (double)ci0);
        const float ci = INV ? -ci0 : ci0, cil = INV ? -cil0 : cil0;
        const float re = fmaf(a.x, cr, fmaf(-a.y, ci, fmaf(a.x, crl, -a.y * cil)));
        const float im = fmaf(a.x, ci, fmaf(a.y, cr, fmaf(a.x, cil, a.y * crl)));
        return mk<T>(re, im);
    }
}

template <typename T, bool INV> __device__ __forceinline__ void dft4(c2<T> &x0, c2<T> &x1, c2<T> &x2, c2<T> &x3) {
    const c2<T> s02 = cadd<T>(x0, x2), d02 = csub<T>(x0, x2);
    const c2<T> s13 = cadd<T>(x1, x3), d13 = mul_wi<T, INV>(csub<T>(x1, x3));
    x0 = cadd<T>(s02, s13);
    x2 = csub<T>(s02, s13);
    x1 = cadd<T>(d02, d13);
    x3 = csub<T>(d02, d13);
}

// r-point DFT in place, natural order in and out
template <typename T, int R, bool INV> struct Dft;

template <typename T, bool INV> struct Dft<T, 2, INV> {
    __device__ __forceinline__ static void run(c2<T> *x) {
        const c2<T> a = x[0];
        x[0] = cadd<T>(a, x[1]);
        x[1] = csub<T>(a, x[1]);
    }
};

template <typename T, bool INV> struct Dft<T, 4, INV> {
    __device__ __forceinline__ static void run(c2<T> *x) {
        dft4<T, INV>(x[0], x[1], x[2], x[3]);
    }
};

// n = 2 n1 + n2, k = k1 + 4 k2: X[k] = sum_n2 W8^{n2 k1} (-1)^{n2 k2} DFT4_n1(x[2 n1 + n2])[k1]
template <typename T, bool INV> struct Dft<T, 8, INV> {
    __device__ __forceinline__ static void run(c2<T> *x) {
        const double h = 0.70710678118654752440;
        dft4<T, INV>(x[0], x[2], x[4], x[6]);
        dft4<T, INV>(x[1], x[3], x[5], x[7]);
        const c2<T> b1 = mul_k<T, INV>(x[3], h, -h);
        const c2<T> b2 = mul_wi<T, INV>(x[5]);
        const c2<T> b3 = mul_k<T, INV>(x[7], -h, -h);
        const c2<T> a0 = x[0], a1 = x[2], a2 = x[4], a3 = x[6], b0 = x[1];
        x[0] = cadd<T>(a0, b0);
        x[4] = csub<T>(a0, b0);
        x[1] = cadd<T>(a1, b1);
        x[5] = csub<T>(a1, b1);
        x[2] = cadd<T>(a2, b2);
        x[6] = csub<T>(a2, b2);
        x[3] = cadd<T>(a3, b3);
        x[7] = csub<T>(a3, b3);
    }
};

// n = 4 n1 + n2, k = k1 + 4 k2: X[k] = DFT4_n2( W16^{n2 k1} DFT4_n1(x[4 n1 + n2])[k1] )[k2]
template <typename T, bool INV> struct Dft<T, 16, INV> {
    __device__ __forceinline__ static void run(c2<T> *x) {
        const double c = 0.92387953251128675613, s = 0.38268343236508977173, h = 0.70710678118654752440;
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) dft4<T, INV>(x[n2], x[4 + n2], x[8 + n2], x[12 + n2]);
        // A[n2][k1] sits at x[n2 + 4 k1]; twiddle W16^{n2 k1}
        x[5] = mul_k<T, INV>(x[5], c, -s);      // 1
        x[9] = mul_k<T, INV>(x[9], h, -h);      // 2
        x[13] = mul_k<T, INV>(x[13], s, -c);    // 3
        x[6] = mul_k<T, INV>(x[6], h, -h);      // 2
        x[10] = mul_wi<T, INV>(x[10]);          // 4
        x[14] = mul_k<T, INV>(x[14], -h, -h);   // 6
        x[7] = mul_k<T, INV>(x[7], s, -c);      // 3
        x[11] = mul_k<T, INV>(x[11], -h, -h);   // 6
        x[15] = mul_k<T, INV>(x[15], -c, s);    // 9
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) dft4<T, INV>(x[4 * k1], x[4 * k1 + 1], x[4 * k1 + 2], x[4 * k1 + 3]);
        // X[k1 + 4 k2] sits at x[4 k1 + k2]: transpose the 4 x 4 index
        c2<T> y[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) y[q] = x[4 * (q & 3) + (q >> 2)];
#pragma unroll
        for (int q = 0; q < 16; ++q) x[q] = y[q];
    }
};

// v[j] *= w^{jk} (conjugated for the inverse) from a table row of r entries
// Twiddle rows of one pass into registers: butterfly b of this thread takes
// w[b R + j] = row (i_b mod p), entry j (P entries per thread in every pass).
template <typename T, int R>
__device__ __forceinline__ void load_rows(c2<T> *w, const c2<T> *__restrict__ tbl, int M, int lp) {
    constexpr int PTS = fft_pts<T>(), NBF = PTS / R;
    const int ntf = M / PTS, p = 1 << lp, t = threadIdx.x;
    if (t < ntf) {
#pragma unroll
        for (int b = 0; b < NBF; ++b) {
            const c2<T> *row = tbl + (size_t)((t + b * ntf) & (p - 1)) * R;
            if constexpr (sizeof(T) == 4) {
                const float4 *w4 = reinterpret_cast<const float4 *>(row);
#pragma unroll
                for (int q = 0; q < R / 2; ++q) {
                    const float4 x = w4[q];
                    w[b * R + 2 * q] = mk<T>(x.x, x.y);
                    w[b * R + 2 * q + 1] = mk<T>(x.z, x.w);
                }
            } else {
#pragma unroll
                for (int j = 0; j < R; ++j) w[b * R + j] = row[j];
            }
        }
    }
}

// One pass.  load(q, pq) and post(k, pk, x) get the natural index and its
// padded LDS index.  Padded addresses advance linearly in j: the read stride
// M/R and the write stride p of every pass after the first are multiples of
// P, and the first pass writes P consecutive elements of one P-block.
// FP32: w holds this pass's twiddle rows (load_rows), already in registers;
// pre() runs after the butterflies and before the barrier and fetches the
// next pass's rows into w, so their latency overlaps the exchange.  FP64
// (twice the registers per row): rows are read from the table tw as they
// are applied.
template <typename T, int R, bool INV, bool FIRST, bool LAST, class Load, class Post, class Pre>
__device__ __forceinline__ void fft_pass(c2<T> *buf, int M, int lp, c2<T> *w, const c2<T> *__restrict__ tw,
                                         Load &load, Post &post, Pre pre) {
    constexpr int LP = fft_lp<T>(), PTS = 1 << LP;
    constexpr int NBF = PTS / R;                        // butterflies per thread
    const bool full = M >= PTS;                         // every FFT thread has NBF butterflies
    const int ntf = full ? M / PTS : 1;
    const int nbv = full ? NBF : 1;
    const int Mr = M / R, p = 1 << lp;
    const int sMr = Mr + (Mr >> LP), sp = FIRST ? 1 : p + (p >> LP);
    const int t = threadIdx.x;
    c2<T> v[NBF][R];
    if (t < ntf) {
#pragma unroll
        for (int b = 0; b < NBF; ++b) {
            if (b < nbv) {
                const int i = t + b * ntf;
                if constexpr (FIRST) {
                    const bool lin = (Mr & (PTS - 1)) == 0;
#pragma unroll
                    for (int j = 0; j < R; ++j) {
                        const int q = i + j * Mr;
                        v[b][j] = load(q, lin ? pad_idx<T>(i) + j * sMr : pad_idx<T>(q));
                    }
                } else {
                    const int pi = pad_idx<T>(i);
#pragma unroll
                    for (int j = 0; j < R; ++j) v[b][j] = buf[pi + j * sMr];
                    if constexpr (sizeof(T) == 4) {
#pragma unroll
                        for (int j = 1; j < R; ++j) v[b][j] = mul_c<T, INV>(v[b][j], w[b * R + j].x, w[b * R + j].y);
                    } else {
                        const c2<T> *row = tw + (size_t)(i & (p - 1)) * R;
#pragma unroll
                        for (int j = 1; j < R; ++j) {
                            const c2<T> x = row[j];
                            v[b][j] = mul_c<T, INV>(v[b][j], x.x, x.y);
                        }
                    }
                }
                Dft<T, R, INV>::run(v[b]);
            }
        }
    }
    pre();
    __syncthreads();
    if (t < ntf) {
#pragma unroll
        for (int b = 0; b < NBF; ++b) {
            if (b < nbv) {
                const int i = t + b * ntf;
                const int k = FIRST ? 0 : (i & (p - 1));
                const int base = (i - k) * R + k;
                const int pb = pad_idx<T>(base);
#pragma unroll
                for (int j = 0; j < R; ++j) {
                    if constexpr (LAST) post(base + j * p, pb + j * sp, v[b][j]);
                    else buf[pb + j * sp] = v[b][j];
                }
            }
        }
    }
    __syncthreads();
}

// Full FFT of size M = 2^log2M (forward e^{-i}, inverse e^{+i},
// unnormalised): input from load(q, pq), q < M; output through
// post(k, pk, X_k); pq, pk: padded LDS indices of q, k.
// Every thread of the workgroup must call it (it synchronises).
template <typename T, bool INV, class Load, class Post>
__device__ __forceinline__ void fft_run(c2<T> *buf, int M, int log2M, const c2<T> *__restrict__ tw, Load load,
                                        Post post) {
    constexpr int LP = fft_lp<T>(), PTS = 1 << LP;
    const int nmain = log2M / LP, rem = log2M % LP;
    c2<T> w[PTS];
    auto none = [] {};
    if (nmain == 0) {
        switch (rem) {
        case 0: {
            c2<T> v = mk<T>(0, 0);
            if (threadIdx.x == 0) v = load(0, 0);
            __syncthreads();
            if (threadIdx.x == 0) post(0, 0, v);
            __syncthreads();
            return;
        }
        case 1: fft_pass<T, 2, INV, true, true>(buf, M, 0, w, tw, load, post, none); return;
        case 2: fft_pass<T, 4, INV, true, true>(buf, M, 0, w, tw, load, post, none); return;
        default:
            if constexpr (PTS > 8) fft_pass<T, 8, INV, true, true>(buf, M, 0, w, tw, load, post, none);
            return;
        }
    }
    if (nmain == 1 && rem == 0) {
        fft_pass<T, PTS, INV, true, true>(buf, M, 0, w, tw, load, post, none);
        return;
    }
    // rows of the pass after the one with stride 2^lp (table at twp)
    auto fetch = [&](int s_next, int lp_next, const c2<T> *twp_next) {
        if (s_next < nmain) {
            load_rows<T, PTS>(w, twp_next, M, lp_next);
        } else {
            switch (rem) {
            case 1: load_rows<T, 2>(w, twp_next, M, lp_next); break;
            case 2: load_rows<T, 4>(w, twp_next, M, lp_next); break;
            case 3:
                if constexpr (PTS > 8) load_rows<T, 8>(w, twp_next, M, lp_next);
                break;
            default: break;
            }
        }
    };
    // FP32: the rows of pass s + 1 are fetched before pass s's barrier
    constexpr bool PF = sizeof(T) == 4;
    const c2<T> *twp = tw;
    fft_pass<T, PTS, INV, true, false>(buf, M, 0, w, twp, load, post, [&] { if (PF) fetch(1, LP, twp); });
    int lp = LP;
    for (int s = 1; s < nmain; ++s) {
        const c2<T> *twn = twp + ((size_t)PTS << lp);
        auto pre = [&] { if (PF && (s + 1 < nmain || rem)) fetch(s + 1, lp + LP, twn); };
        if (s == nmain - 1 && rem == 0) fft_pass<T, PTS, INV, false, true>(buf, M, lp, w, twp, load, post, none);
        else fft_pass<T, PTS, INV, false, false>(buf, M, lp, w, twp, load, post, pre);
        twp = twn;
        lp += LP;
    }
    switch (rem) {
    case 1: fft_pass<T, 2, INV, false, true>(buf, M, lp, w, twp, load, post, none); break;
    case 2: fft_pass<T, 4, INV, false, true>(buf, M, lp, w, twp, load, post, none); break;
    case 3:
        if constexpr (PTS > 8) fft_pass<T, 8, INV, false, true>(buf, M, lp, w, twp, load, post, none);
        break;
    default: break;
    }
}

// entries of the twiddle table of class M for 2^LP points per thread (same
// pass walk as fft_run)
inline int64_t fft_table_entries(int M, int LP) {
    const int L = ilog2(M), nmain = L / LP, rem = L % LP;
    int64_t e = 0;
    int lp = LP;
    for (int s = 1; s < nmain; ++s, lp += LP) e += (int64_t)1 << (LP + lp);
    if (nmain >= 1 && rem) e += (int64_t)(1 << rem) << lp;
    return e;
}

// fill the table of class M (2^LP points per thread), rounded from long double
template <typename C> inline void fft_table_fill(int M, int LP, C *t) {
    const long double pi = 3.141592653589793238462643383279502884L;
    const int L = ilog2(M), nmain = L / LP, rem = L % LP;
    int64_t o = 0;
    auto pass = [&](int r, int lp) {
        const int64_t p = (int64_t)1 << lp, rp = r * p;
        for (int64_t k = 0; k < p; ++k)
            for (int j = 0; j < r; ++j) {
                const int64_t e = (j * k) % rp;
                const long double a = -2.0L * pi * (long double)e / (long double)rp;
                t[o].x = cosl(a);
                t[o].y = sinl(a);
                ++o;
            }
    };
    int lp = LP;
    for (int s = 1; s < nmain; ++s, lp += LP) pass(1 << LP, lp);
    if (nmain >= 1 && rem) pass(1 << rem, lp);
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

// launch bound: FP32 classes stop at M = 16384, FP64 at M = 8192 (LDS)
template <typename T> constexpr int ring_max_threads() { return 1024; }

// Block g of a ring launch -> (ring slot, map): the maps of one ring run
// back to back on one XCD (consecutive blocks go round-robin over the 8
// XCDs), so its chirp spectrum, chirp and twiddle rows are read into that
// XCD's L2 once.  Grid: ceil(ncls / 8) * 8 * nmaps blocks.
__device__ __forceinline__ void ring_block(int g, int nmaps, int *slot, int *map) {
    const int xcd = g & 7, k = g >> 3;
    *map = k % nmaps;
    *slot = (k / nmaps) * 8 + xcd;
}
inline unsigned ring_grid(int ncls, int nmaps) { return (unsigned)(((ncls + 7) / 8) * 8 * nmaps); }
// synthesis: each XCD takes a contiguous run of slots (octets q of C per
// XCD), the 8 slots of an octet next to each other in time, so that the
// 64-byte lines of the synthesis layout (8 ring pairs of one m, shared by 8
// consecutive rings of a class) meet in that XCD's L2.  Bluestein classes
// run octet-major, the maps of one octet back to back, so the octet's chirp
// spectra and chirps (up to 8 x 96 KB) stay in L2 from one map to the next
// (FETCH per 32 maps at nside 2048, 8192-point class: 3.7 GB against 9.7
// map-major, for 2.1 GB of coefficients); power-of-two classes, which have
// no chirps, run map-major (4.4 GB against 6.7 octet-major, for 4.3 GB)
template <bool BLUE>
__device__ __forceinline__ void ring_block_syn(int g, int nmaps, int C, int *slot, int *map) {
    const int xcd = g & 7, k = g >> 3;
    if constexpr (BLUE) {
        *map = (k >> 3) % nmaps;
        *slot = ((xcd * C + (k >> 3) / nmaps) << 3) + (k & 7);
    } else {
        *map = (k >> 3) / C;
        *slot = ((xcd * C + (k >> 3) % C) << 3) + (k & 7);
    }
}
inline unsigned ring_grid_syn(int ncls, int nmaps) { return (unsigned)(((ncls + 63) / 64) * 64 * nmaps); }

// ---------------------------------------------------------------------------
// analysis: maps -> phase.  BLUE: Bluestein rings (h not a power of two),
// V + vofs[r]: chirp spectrum (M entries) then the chirp c_j, j < h.
// TP: precision of the phase written (FP64 transforms of FP32 spin plans
// round their output once to FP32).
// ---------------------------------------------------------------------------
// One ring r of one map: mapb / phb are that map's pixels and phase buffer.
// Run by the fused ring launches (below), for one plan or for every group of
// a mixed-geometry batch.
template <typename T, typename TIn, bool BLUE, typename TP>
__device__ __forceinline__ void ring_analysis_ring(c2<T> *buf, const TIn *__restrict__ mapb,
                                                   c2<TP> *__restrict__ phb, int r,
                                                   const int64_t *__restrict__ ring_start,
                                                   const int32_t *__restrict__ nphi,
                                                   const int32_t *__restrict__ mcut, const double *__restrict__ phi0,
                                                   const double *__restrict__ scale,
                                                   const double2 *__restrict__ rot, const c2<T> *__restrict__ tw,
                                                   const c2<T> *__restrict__ V, const int64_t *__restrict__ vofs,
                                                   int M, int log2M, int mmax, int nr) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int n = nphi[r], h = n >> 1;
    const TIn *f = mapb + ring_start[r];
    auto pixel_pair = [&](int q) { return mk<T>(load_as<T>(f + 2 * q), load_as<T>(f + 2 * q + 1)); };
    auto lds = [&](int, int pq) { return buf[pq]; };
    if constexpr (!BLUE) {
        fft_run<T, false>(
            buf, M, log2M, tw, [&](int q, int) { return pixel_pair(q); },
            [&](int, int pk, c2<T> x) { buf[pk] = x; });
    } else {
        const c2<T> *Vr = V + vofs[r];
        const c2<T> *Cr = Vr + M;
        fft_run<T, false>(
            buf, M, log2M, tw,
            [&](int q, int) {
                c2<T> z = mk<T>(0, 0);
                if (q < h) z = cmul<T>(pixel_pair(q), cconj<T>(Cr[q]));
                return z;
            },
            [&](int k, int pk, c2<T> x) { buf[pk] = cmul<T>(x, Vr[k]); });
        const T invM = (T)1 / (T)M;
        fft_run<T, true>(buf, M, log2M, tw, lds, [&](int k, int pk, c2<T> x) {
            if (k < h) buf[pk] = cscale<T>(cmul<T>(x, cconj<T>(Cr[k])), invM);
        });
    }
    // untangle in place: Z (period h) -> X[k] = E_k + W_k O_k for k = 0..h,
    //   E_k = (Z[k] + conj Z[-k]) / 2,  O_k = (Z[k] - conj Z[-k]) / 2i,
    //   W_k = e^{-2 pi i k / n};
    // thread k <= h/2 overwrites the two slots it reads, k and h-k, with X[k]
    // and X[h-k] = conj(E_k - W_k O_k) (k = 0: X[h] into slot h, one past the
    // FFT).  The rest of the spectrum is X[n-k] = conj X[k] (real input).
    {
        double2 W;
        double s, c;
        sincospi(-2.0 * (double)(tid % n) / (double)n, &s, &c);
        W = make_double2(c, s);
        const double2 Wst = rot[3 * nr + r];
        for (int k = tid; 2 * k <= h; k += nt) {
            const c2<T> za = buf[pad_idx<T>(k)];
            const c2<T> zb = cconj<T>(buf[pad_idx<T>(k == 0 ? 0 : h - k)]);
            const c2<T> E = cscale<T>(cadd<T>(za, zb), (T)0.5);
            const c2<T> O = cmul_mi<T>(cscale<T>(csub<T>(za, zb), (T)0.5));
            const c2<T> Tk = cmul<T>(round_c<T>(W), O);
            buf[pad_idx<T>(k)] = cadd<T>(E, Tk);
            if (2 * k != h) buf[pad_idx<T>(h - k)] = cconj<T>(csub<T>(E, Tk));
            W = zmul(W, Wst);
        }
    }
    __syncthreads();
    // F(m) = [m < mcut] scale e^{-i m phi0} X[m mod n], m = 0..mmax: thread
    // tid takes m = tid + j nt (nt a multiple of 8, so its stores step by a
    // fixed number of m-blocks; lanes fill 64-byte m-blocks), the rotation
    // times the scale by FP64 steps e^{-i nt phi0} from one sincos, rounded
    // once per output.
    const double scd = scale ? scale[r] : 1.0;
    const int cut = mcut[r];
    const double ph0 = phi0 ? phi0[r] : 0.0;
    const bool rotp = (ph0 != 0.0);
    const T sc = (T)scd;
    c2<TP> *o = phb + ((int64_t)(tid >> 3) * nr + r) * MB + (tid & 7);
    const int64_t ostep = (int64_t)(nt >> 3) * nr * MB;
    double2 P = make_double2(scd, 0.0);
    if (rotp) {
        double s, c;
        sincos(-(double)tid * ph0, &s, &c);
        P = make_double2(scd * c, scd * s);
    }
    const double2 Pst = zconj(rot[r]);                  // e^{-i nt phi0}
    const int mstep = nt % n;
    int mm = tid % n;                                   // m mod n
    for (int m = tid; m <= mmax; m += nt, o += ostep) {
        c2<T> F = mk<T>(0, 0);
        if (m < cut) {
            const bool low = mm <= h;
            c2<T> X = buf[pad_idx<T>(low ? mm : n - mm)];
            if (!low) X.y = -X.y;
            F = rotp ? cmul<T>(X, round_c<T>(P)) : cscale<T>(X, sc);
        }
        *o = mk<TP>((TP)F.x, (TP)F.y);
        if (rotp) P = zmul(P, Pst);
        mm += mstep;
        if (mm >= n) mm -= n;
    }
}

// ---------------------------------------------------------------------------
// synthesis: phase -> maps
// ---------------------------------------------------------------------------
// One ring r of one map: phb / mapb are that map's phase buffer and pixels.
template <typename T, typename TOut, bool BLUE>
__device__ __forceinline__ void ring_synthesis_ring(c2<T> *buf, const c2<T> *__restrict__ phb,
                                                    TOut *__restrict__ mapb, int r,
                                                    const int64_t *__restrict__ ring_start,
                                                    const int32_t *__restrict__ nphi,
                                                    const int32_t *__restrict__ mcut,
                                                    const double *__restrict__ phi0,
                                                    const double *__restrict__ scale,
                                                    const double2 *__restrict__ rot, const c2<T> *__restrict__ tw,
                                                    const c2<T> *__restrict__ V, const int64_t *__restrict__ vofs,
                                                    int M, int log2M, int mmax, int nr) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int n = nphi[r], h = n >> 1;
    const int top = min(mmax, mcut[r] - 1);
    const c2<T> *Vr = BLUE ? V + vofs[r] : nullptr;
    const c2<T> *Cr = BLUE ? Vr + M : nullptr;
    const double ph0 = phi0 ? phi0[r] : 0.0;
    const bool rotp = (ph0 != 0.0);
    // synthesis layout: ring r is half (r >= npairs) of ring pair p
    const int npairs = (nr + 1) >> 1;
    const int half = r >= npairs ? 1 : 0, pr = half ? nr - 1 - r : r;
    const int64_t sbe = syn_block_elems(npairs);
    const c2<T> *G = phb + syn_ofs(0, half, pr, npairs);
    const double2 Rn = rot[nr + r];
    auto gload = [&](int m) { return G[(int64_t)(m >> 3) * sbe + (m & 7) * 8]; };

    // X[k] = sum_{m = k mod n} G'(m) + sum_{m = -k mod n, m >= 1} conj G'(m),
    // G'(m) = G(m) e^{i m phi0} (imaginary part of G(0) dropped); Rk = e^{i k phi0}
    auto xsum = [&](int k, double2 Rk) -> c2<T> {
        c2<T> acc = mk<T>(0, 0);
        double2 q = Rk;
        for (int m = k; m <= top; m += n) {
            c2<T> g = gload(m);
            if (m == 0) g.y = 0;
            if (rotp) {
                g = cmul<T>(g, round_c<T>(q));
                q = zmul(q, Rn);
            }
            acc = cadd<T>(acc, g);
        }
        q = zmul(zconj(Rk), Rn);
        for (int m = (k == 0 ? n : n - k); m <= top; m += n) {
            c2<T> g = gload(m);
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
    // top < n (no alias copies): bins k and h - k take m in {k, n-k} and {h-k, h+k}
    const bool direct = top < n;
    for (int k = tid; k <= (h >> 1); k += nt) {
        const int k2 = h - k;
        c2<T> Xa, Xb;
        if (direct) {
            const c2<T> z0 = mk<T>(0, 0);
            c2<T> g0 = k <= top ? gload(k) : z0;
            c2<T> g1 = (k >= 1 && n - k <= top) ? gload(n - k) : z0;
            c2<T> g2 = k2 <= top ? gload(k2) : z0;
            c2<T> g3 = h + k <= top ? gload(h + k) : z0;
            if (k == 0) g0.y = 0;
            if (rotp) {
                // R_{n-k} = R_n conj R_k, R_{h-k} = R_h conj R_k, R_{h+k} = R_h R_k
                g0 = cmul<T>(g0, round_c<T>(R));
                g1 = cmul<T>(g1, round_c<T>(zmul(Rn, zconj(R))));
                g2 = cmul<T>(g2, round_c<T>(zmul(Rh, zconj(R))));
                g3 = cmul<T>(g3, round_c<T>(zmul(Rh, R)));
            }
            Xa = cadd<T>(g0, cconj<T>(g1));
            Xb = cadd<T>(g2, cconj<T>(g3));
        } else {
            Xa = xsum(k, R);
            Xb = xsum(k2, rotp ? zmul(Rh, zconj(R)) : R);
        }
        const c2<T> Wk = round_c<T>(W);
        const c2<T> Wk2 = mk<T>(-Wk.x, Wk.y);        // e^{2 pi i (h - k)/n} = -conj W_k
        {
            const c2<T> s = cadd<T>(Xa, cconj<T>(Xb)), d = csub<T>(Xa, cconj<T>(Xb));
            c2<T> z = cadd<T>(s, cmul_pi<T>(cmul<T>(Wk, d)));
            if (BLUE) z = cconj<T>(cmul<T>(z, Cr[k]));
            buf[pad_idx<T>(k)] = z;
        }
        if (k != 0 && k2 != k) {
            const c2<T> s = cadd<T>(Xb, cconj<T>(Xa)), d = csub<T>(Xb, cconj<T>(Xa));
            c2<T> z = cadd<T>(s, cmul_pi<T>(cmul<T>(Wk2, d)));
            if (BLUE) z = cconj<T>(cmul<T>(z, Cr[k2]));
            buf[pad_idx<T>(k2)] = z;
        }
        W = zmul(W, Wst);
        if (rotp) R = zmul(R, S);
    }
    if (BLUE)
        for (int j = h + tid; j < M; j += nt) buf[pad_idx<T>(j)] = mk<T>(0, 0);
    __syncthreads();
    const T sc = (T)(scale ? scale[r] : 1.0);
    TOut *f = mapb + ring_start[r];
    auto lds = [&](int, int pq) { return buf[pq]; };
    if constexpr (!BLUE) {
        fft_run<T, true>(buf, M, log2M, tw, lds, [&](int j, int, c2<T> z) {
            f[2 * j] = (TOut)(z.x * sc);
            f[2 * j + 1] = (TOut)(z.y * sc);
        });
    } else {
        fft_run<T, false>(buf, M, log2M, tw, lds, [&](int k, int pk, c2<T> x) { buf[pk] = cmul<T>(x, Vr[k]); });
        const T invM = (T)1 / (T)M;
        fft_run<T, true>(buf, M, log2M, tw, lds, [&](int j, int, c2<T> x) {
            if (j < h) {
                const c2<T> z = cscale<T>(cmul<T>(Cr[j], cconj<T>(x)), invM);
                f[2 * j] = (TOut)(z.x * sc);
                f[2 * j + 1] = (TOut)(z.y * sc);
            }
        });
    }
}

// ---------------------------------------------------------------------------
// FP64 low-m band of FP32 plans: direct DFT in FP64 for m < M_LO,
//   F_r(m) = [m < mcut_r] scale_r e^{-i m phi0_r} sum_j f_j e^{-2 pi i m j / n}
// (same conventions as the ring analysis), on the FP64 matrix cores: per
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

// The band for batches below the matrix-core threshold (single maps): one
// wave per (ring, map) on the VALU, each lane summing its pixels
// j = lane + 64 t against e^{-2 pi i m j / n}, m < M_LO (e^{-2 pi i j / n}
// from one sincospi per lane, advanced to the lane's next pixel by an FP64
// rotation step: at most n / 64 <= 256 steps, a relative error of order
// 1e-14; the higher m by complex multiplication), then a fixed-order xor
// reduction over the wave, after which every lane holds the same sums.  Same
// conventions and output as ring_band_kernel.
constexpr int RBV_LANES = 64;

// One ring r of one map (mapb: its pixels, bandb: its band phase); every
// lane of a one-wave workgroup calls it.
template <typename TIn>
__device__ __forceinline__ void ring_band_valu_ring(const TIn *__restrict__ mapb, double2 *__restrict__ bandb, int r,
                                                    const int64_t *__restrict__ ring_start,
                                                    const int32_t *__restrict__ nphi,
                                                    const int32_t *__restrict__ mcut, const double *__restrict__ phi0,
                                                    const double *__restrict__ scale, int mtop, int nr) {
    const int lane = threadIdx.x;
    const int n = nphi[r];
    const TIn *f = mapb + ring_start[r];
    double acc[2 * M_LO];
#pragma unroll
    for (int i = 0; i < 2 * M_LO; ++i) acc[i] = 0.0;
    double s1, c1, ss, cs;                          // e^{-2 pi i j / n}, j = lane; the step to j + 64
    sincospi(-2.0 * (double)(lane % n) / (double)n, &s1, &c1);
    sincospi(-2.0 * (double)(RBV_LANES % n) / (double)n, &ss, &cs);
    for (int j = lane; j < n; j += RBV_LANES) {
        const double v = (double)f[j];
        double cr = 1.0, ci = 0.0;
#pragma unroll
        for (int m = 0; m < M_LO; ++m) {
            acc[2 * m] = fma(v, cr, acc[2 * m]);
            acc[2 * m + 1] = fma(v, ci, acc[2 * m + 1]);
            if (m + 1 < M_LO) {
                const double t = cr * c1 - ci * s1;
                ci = cr * s1 + ci * c1;
                cr = t;
            }
        }
        const double t = c1 * cs - s1 * ss;
        s1 = c1 * ss + s1 * cs;
        c1 = t;
    }
#pragma unroll
    for (int i = 0; i < 2 * M_LO; ++i) {
        double v = acc[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        acc[i] = v;
    }
    if (lane <= mtop) {
        const int mm = lane;
        double x = 0.0, y = 0.0;
#pragma unroll
        for (int i = 0; i < M_LO; ++i)
            if (i == mm) {
                x = acc[2 * i];
                y = acc[2 * i + 1];
            }
        double2 F = make_double2(0.0, 0.0);
        if (mm < mcut[r]) {
            const double sc = scale ? scale[r] : 1.0;
            x *= sc;
            y *= sc;
            const double ph = phi0 ? phi0[r] : 0.0;
            if (ph != 0.0) {
                double sn, csn;
                sincos(-(double)mm * ph, &sn, &csn);
                F = make_double2(x * csn - y * sn, x * sn + y * csn);
            } else {
                F = make_double2(x, y);
            }
        }
        bandb[((int64_t)(mm >> 3) * nr + r) * MB + (mm & 7)] = F;
    }
}

template <typename TIn>
__global__ void __launch_bounds__(RBV_LANES)
ring_band_valu_kernel(const TIn *__restrict__ maps, int64_t map_stride, double2 *__restrict__ band,
                      int64_t band_stride, const int64_t *__restrict__ ring_start, const int32_t *__restrict__ nphi,
                      const int32_t *__restrict__ mcut, const double *__restrict__ phi0,
                      const double *__restrict__ scale, int mtop, int nr) {
    const int r = blockIdx.x, b = blockIdx.y;
    ring_band_valu_ring<TIn>(maps + (int64_t)b * map_stride, band + (int64_t)b * band_stride, r, ring_start, nphi,
                             mcut, phi0, scale, mtop, nr);
}

int launch_ring_band(const Plan *p, int nmaps, const void *maps, int map_dtype, int64_t map_stride, void *band,
                     const double *scale, hipStream_t s) {
    const int mtop = std::min(M_LO - 1, p->mmax);
    if (nmaps < mfma_min_maps(p)) {
        dim3 grid(p->nrings, nmaps);
        if (map_dtype == HBX_DT_F64)
            hipLaunchKernelGGL(ring_band_valu_kernel<double>, grid, dim3(RBV_LANES), 0, s, (const double *)maps,
                               map_stride, (double2 *)band, band_phase_elems(p), p->d_ring_start, p->d_nphi,
                               p->d_mcut, p->d_phi0, scale, mtop, p->nrings);
        else
            hipLaunchKernelGGL(ring_band_valu_kernel<float>, grid, dim3(RBV_LANES), 0, s, (const float *)maps,
                               map_stride, (double2 *)band, band_phase_elems(p), p->d_ring_start, p->d_nphi,
                               p->d_mcut, p->d_phi0, scale, mtop, p->nrings);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_error(e, "ring_band_valu_kernel launch");
        return 0;
    }
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
__global__ void round_to_f32_kernel(const double2 *__restrict__ src, float2 *__restrict__ dst, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = make_float2((float)src[i].x, (float)src[i].y);
}

// V_h = FFT_M(v), v_j = c_j (0 <= j < h), v_{M-j} = c_j (1 <= j < h), else 0,
// stored at V + ofs followed by the chirp c_j itself (j < h)
template <typename T>
__global__ void __launch_bounds__(ring_max_threads<T>())
chirp_spectrum_kernel(const int32_t *__restrict__ hs, const int64_t *__restrict__ ofs, c2<T> *V,
                      const c2<T> *__restrict__ tw, int M, int log2M) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    c2<T> *buf = reinterpret_cast<c2<T> *>(smem_raw);
    const int h = hs[blockIdx.x];
    c2<T> *out = V + ofs[blockIdx.x];
    for (int j = threadIdx.x; j < h; j += blockDim.x) out[M + j] = chirp<T>(j, h);
    fft_run<T, false>(
        buf, M, log2M, tw,
        [&](int j, int) {
            c2<T> v = mk<T>(0, 0);
            if (j < h) v = chirp<T>(j, h);
            else if (j > M - h) v = chirp<T>(M - j, h);
            return v;
        },
        [&](int k, int, c2<T> x) { out[k] = x; });
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
static int build_chirp_spectra(const std::vector<int32_t> &hs_host, const std::vector<int64_t> &ofs_host, int M,
                               c2<T> *V, const c2<T> *tw) {
    if (hs_host.empty()) return 0;
    int32_t *d_hs = nullptr;
    int64_t *d_ofs = nullptr;
    HBX_HIP(hipMalloc(&d_hs, hs_host.size() * sizeof(int32_t)));
    HBX_HIP(hipMalloc(&d_ofs, ofs_host.size() * sizeof(int64_t)));
    HBX_HIP(hipMemcpy(d_hs, hs_host.data(), hs_host.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    HBX_HIP(hipMemcpy(d_ofs, ofs_host.data(), ofs_host.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    const size_t lds = ring_lds_bytes<T>(M);
    int rc = allow_lds<T>(chirp_spectrum_kernel<T>, lds);
    if (rc == 0) {
        hipLaunchKernelGGL(chirp_spectrum_kernel<T>, dim3((unsigned)hs_host.size()), dim3(class_threads<T>(M)), lds, 0,
                           d_hs, d_ofs, V, tw, M, ilog2(M));
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = hip_error(e, "chirp_spectrum_kernel");
    }
    hipError_t e = hipDeviceSynchronize();
    (void)hipFree(d_hs);
    (void)hipFree(d_ofs);
    if (rc) return rc;
    if (e != hipSuccess) return hip_error(e, "chirp_spectrum_kernel sync");
    return 0;
}

// ---------------------------------------------------------------------------
// Fused ring launches: one launch runs every size class of one workgroup size
// (FP32: 64 threads for M <= 1024, then 128 .. 1024; power-of-two and
// Bluestein rings apart), for one plan or for every group of a mixed-geometry
// batch.  A segment is one plan's rings of one class; a block finds its
// segment by a binary search over the segments' first blocks (scalar loads),
// then maps to (ring slot, map) as a launch of that class alone would: the
// maps of one ring run back to back on one XCD (analysis), and each XCD
// takes a contiguous run of ring octets (synthesis).  Segments start at
// multiples of 8 (analysis) or 64 (synthesis) blocks, so this grouping
// holds in every segment.  The rings run ring_analysis_ring /
// ring_synthesis_ring with the segment's tables; table pointers read from a
// segment are cast to the global address space so their loads and stores
// stay global ones.
// ---------------------------------------------------------------------------
template <typename X> __device__ __forceinline__ X *gptr(X *p) { return (X *)(HBX_GLOBAL X *)p; }

// last segment whose first block (block0 * unit) is <= blk
__device__ __forceinline__ const RingSeg *find_seg(const RingSeg *segs, int nseg, int blk, int unit) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ldc(&segs[mid].block0) * unit <= blk) lo = mid;
        else hi = mid - 1;
    }
    return segs + lo;
}

// Where a block's map and phase buffer are: a batch (maps_tab != null) reads
// them per group from device tables and the workspace; one plan takes the
// call's arguments.
template <typename TM, typename TPh>
__device__ __forceinline__ void seg_buffers(const RingCall &c, const RingSeg *sg, int b, TM *&map, TPh *&ph,
                                            const double *&scale) {
    const int64_t pst = ldc(&sg->pstride);
    if (c.maps_tab != nullptr) {
        const int g = ldc(&sg->g);
        map = gptr((TM *)ldc(c.maps_tab + g)) + (int64_t)b * ldc(c.mstride_tab + g);
        ph = gptr((TPh *)(c.ws + ldc(&sg->phase_off))) + (int64_t)b * pst;
        scale = gptr(ldc(&sg->scale));
    } else {
        map = gptr((TM *)c.maps) + (int64_t)b * c.mstride;
        ph = gptr((TPh *)c.phase) + (int64_t)b * pst;
        scale = gptr(c.scale);
    }
}

template <typename T, typename TIn, bool BLUE, typename TP>
__global__ void __launch_bounds__(ring_max_threads<T>())
fused_ring_analysis_kernel(const RingSeg *__restrict__ segs, int nseg, RingCall c) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int blk = blockIdx.x;
    const RingSeg *sg = find_seg(segs, nseg, blk, c.unit);
    const int nmaps = c.maps_tab != nullptr ? ldc(&sg->nmaps) : c.nmaps;
    int slot, b;
    ring_block(blk - ldc(&sg->block0) * c.unit, nmaps, &slot, &b);
    if (slot >= ldc(&sg->ncls)) return;
    const TIn *mapb;
    c2<TP> *phb;
    const double *scale;
    seg_buffers<const TIn, c2<TP>>(c, sg, b, mapb, phb, scale);
    ring_analysis_ring<T, TIn, BLUE, TP>(
        reinterpret_cast<c2<T> *>(smem_raw), mapb, phb, ldc(ldc(&sg->rings) + slot), gptr(ldc(&sg->ring_start)),
        gptr(ldc(&sg->nphi)), gptr(ldc(&sg->mcut)), gptr(ldc(&sg->phi0)), scale, gptr(ldc(&sg->rot)),
        gptr((const c2<T> *)ldc(&sg->tw)), gptr((const c2<T> *)ldc(&sg->V)), gptr(ldc(&sg->vofs)), ldc(&sg->M),
        ldc(&sg->log2M), ldc(&sg->mmax), ldc(&sg->nr));
}

template <typename T, typename TOut, bool BLUE>
__global__ void __launch_bounds__(ring_max_threads<T>())
fused_ring_synthesis_kernel(const RingSeg *__restrict__ segs, int nseg, RingCall c) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int blk = blockIdx.x;
    const RingSeg *sg = find_seg(segs, nseg, blk, c.unit);
    const int nmaps = c.maps_tab != nullptr ? ldc(&sg->nmaps) : c.nmaps;
    const int ncls = ldc(&sg->ncls);
    int slot, b;
    ring_block_syn<BLUE>(blk - ldc(&sg->block0) * c.unit, nmaps, (ncls + 63) >> 6, &slot, &b);
    if (slot >= ncls) return;
    TOut *mapb;
    const c2<T> *phb;
    const double *scale;
    seg_buffers<TOut, const c2<T>>(c, sg, b, mapb, phb, scale);
    ring_synthesis_ring<T, TOut, BLUE>(
        reinterpret_cast<c2<T> *>(smem_raw), phb, mapb, ldc(ldc(&sg->rings) + slot), gptr(ldc(&sg->ring_start)),
        gptr(ldc(&sg->nphi)), gptr(ldc(&sg->mcut)), gptr(ldc(&sg->phi0)), scale, gptr(ldc(&sg->rot)),
        gptr((const c2<T> *)ldc(&sg->tw)), gptr((const c2<T> *)ldc(&sg->V)), gptr(ldc(&sg->vofs)), ldc(&sg->M),
        ldc(&sg->log2M), ldc(&sg->mmax), ldc(&sg->nr));
}

// band DFT of a batch (FP32): segment = one group, blocks (map, ring) map-major
template <typename TIn>
__global__ void __launch_bounds__(RBV_LANES)
multi_ring_band_kernel(const RingSeg *__restrict__ segs, int nseg, RingCall c) {
    const int blk = blockIdx.x;
    const RingSeg *sg = find_seg(segs, nseg, blk, 1);
    const int nr = ldc(&sg->nr), local = blk - ldc(&sg->block0);
    const int b = local / nr, r = local - b * nr;
    if (b >= ldc(&sg->nmaps)) return;
    const TIn *mapb;
    double2 *bandb;
    const double *scale;
    seg_buffers<const TIn, double2>(c, sg, b, mapb, bandb, scale);
    const int mmax = ldc(&sg->mmax);
    ring_band_valu_ring<TIn>(mapb, bandb, r, gptr(ldc(&sg->ring_start)), gptr(ldc(&sg->nphi)),
                             gptr(ldc(&sg->mcut)), gptr(ldc(&sg->phi0)), scale, mmax < M_LO - 1 ? mmax : M_LO - 1,
                             nr);
}

template <typename T, typename TP> static int allow_fused_lds(size_t lds) {
    int rc;
    if ((rc = allow_lds<T>(fused_ring_analysis_kernel<T, float, false, TP>, lds))) return rc;
    if ((rc = allow_lds<T>(fused_ring_analysis_kernel<T, double, false, TP>, lds))) return rc;
    if ((rc = allow_lds<T>(fused_ring_analysis_kernel<T, float, true, TP>, lds))) return rc;
    if ((rc = allow_lds<T>(fused_ring_analysis_kernel<T, double, true, TP>, lds))) return rc;
    if constexpr (sizeof(T) == sizeof(TP)) {
        if ((rc = allow_lds<T>(fused_ring_synthesis_kernel<T, float, false>, lds))) return rc;
        if ((rc = allow_lds<T>(fused_ring_synthesis_kernel<T, double, false>, lds))) return rc;
        if ((rc = allow_lds<T>(fused_ring_synthesis_kernel<T, float, true>, lds))) return rc;
        if ((rc = allow_lds<T>(fused_ring_synthesis_kernel<T, double, true>, lds))) return rc;
    }
    return 0;
}

// one class (kind blue) of plan p as a segment; t64: the FP64 tables (FP64
// plans, or the FP64 analysis transforms of an FP32 plan)
static RingSeg class_seg(const Plan *p, const FftClass &c, int blue, bool t64, const double2 *rot) {
    RingSeg sg{};
    sg.rings = blue ? c.d_rings_blue : c.d_rings_pow2;
    sg.ring_start = p->d_ring_start;
    sg.nphi = p->d_nphi;
    sg.mcut = p->d_mcut;
    sg.phi0 = p->d_phi0;
    sg.scale = p->d_scale_ana;
    sg.rot = rot;
    sg.tw = t64 ? (const void *)(p->d_ftw64 + c.tw_off64) : (const void *)(p->d_ftw32 + c.tw_off32);
    sg.V = t64 ? (const void *)p->d_V64 : (const void *)p->d_V32;
    sg.vofs = p->d_vofs;
    sg.pstride = phase_elems_per_map(p);
    sg.ncls = blue ? c.nrings_blue : c.nrings_pow2;
    sg.M = c.M;
    sg.log2M = c.log2M;
    sg.mmax = p->mmax;
    sg.nr = p->nrings;
    return sg;
}

// Launches from segments keyed by (workgroup size, Bluestein), largest
// workgroups first.  absolute: block0 / the grids count blocks (batches,
// each segment with its own map count); otherwise they count units of 8
// (analysis) or 64 (synthesis) blocks per map of the call (one plan).
// alloc(bytes) returns device memory owned by the caller's object.
template <class Alloc>
static int make_launches(const std::map<std::pair<int, int>, std::vector<RingSeg>> &by, bool t64, bool tp32,
                         bool absolute, std::vector<RingLaunch> &out, Alloc alloc) {
    out.clear();
    const int64_t lim = INT32_MAX;
    for (auto it = by.rbegin(); it != by.rend(); ++it) {
        RingLaunch L;
        L.threads = it->first.first;
        L.blue = it->first.second;
        std::vector<RingSeg> ana = it->second, syn = it->second;
        int64_t ba = 0, bs = 0;
        for (size_t i = 0; i < ana.size(); ++i) {
            ana[i].block0 = (int32_t)ba;
            syn[i].block0 = (int32_t)bs;
            syn[i].scale = nullptr;                 // batch synthesis: unit scale (one plan: the call's)
            const int n = ana[i].ncls;
            ba += absolute ? (int64_t)ring_grid(n, ana[i].nmaps) : (n + 7) / 8;
            bs += absolute ? (int64_t)ring_grid_syn(n, syn[i].nmaps) : (n + 63) / 64;
            if (ba > lim || bs > lim) return set_error(HBX_EUNSUPPORTED, "too many rings for one ring launch");
            L.lds = std::max(L.lds, t64 ? ring_lds_bytes<double>(ana[i].M) : ring_lds_bytes<float>(ana[i].M));
        }
        L.nseg = (int)ana.size();
        L.size_ana = ba;
        L.size_syn = bs;
        void *da = alloc(ana.size() * sizeof(RingSeg)), *ds = alloc(syn.size() * sizeof(RingSeg));
        if (!da || !ds) return set_error(HBX_ENOMEM, "hipMalloc for the ring launch tables failed");
        HBX_HIP(hipMemcpy(da, ana.data(), ana.size() * sizeof(RingSeg), hipMemcpyHostToDevice));
        HBX_HIP(hipMemcpy(ds, syn.data(), syn.size() * sizeof(RingSeg), hipMemcpyHostToDevice));
        L.d_ana = static_cast<RingSeg *>(da);
        L.d_syn = static_cast<RingSeg *>(ds);
        int rc = t64 ? (tp32 ? allow_fused_lds<double, float>(L.lds) : allow_fused_lds<double, double>(L.lds))
                     : allow_fused_lds<float, float>(L.lds);
        if (rc) return rc;
        out.push_back(L);
    }
    return HBX_OK;
}

// the plan's own launches: its precision, and (FP32 plans whose analysis
// runs in FP64) the FP64 analysis launches
static int plan_build_fused(Plan *p) {
    auto alloc = [p](size_t bytes) -> void * {
        void *ptr = nullptr;
        if (hipMalloc(&ptr, bytes) != hipSuccess) return nullptr;
        p->allocations.push_back(ptr);
        p->device_bytes += bytes;
        return ptr;
    };
    const bool f64 = p->precision == HBX_F64;
    std::map<std::pair<int, int>, std::vector<RingSeg>> by, by64;
    for (const FftClass &c : p->classes)
        for (int blue = 0; blue < 2; ++blue) {
            if ((blue ? c.nrings_blue : c.nrings_pow2) == 0) continue;
            by[{c.threads, blue}].push_back(class_seg(p, c, blue, f64, p->d_rot));
            if (p->ring_f64) by64[{class_threads<double>(c.M), blue}].push_back(class_seg(p, c, blue, true, p->d_rot64));
        }
    int rc = make_launches(by, f64, false, false, p->fused, alloc);
    if (!rc && p->ring_f64) rc = make_launches(by64, true, true, false, p->fused64, alloc);
    return rc;
}

int plan_build_fft_tables(Plan *p, const std::vector<int32_t> &nphi, const double *phi0) {
    const int nr = p->nrings;
    // class per ring
    std::map<int, std::vector<int32_t>> pow2_rings, blue_rings;   // M -> rings
    std::map<int, int64_t> vofs_of_h;                   // h -> offset (non-pow2)
    std::map<int, std::vector<int32_t>> hs_by_M;        // M -> distinct h
    std::map<int, std::vector<int64_t>> ofs_by_M;
    std::vector<int64_t> vofs(nr, -1);
    std::vector<int> ringM(nr);
    int maxM = 1;
    int64_t vtotal = 0;
    for (int r = 0; r < nr; ++r) {
        const int h = nphi[r] / 2;
        const bool pow2 = (h & (h - 1)) == 0;
        const int M = pow2 ? h : (1 << ilog2(2 * (int64_t)h - 1));
        ringM[r] = M;
        (pow2 ? pow2_rings : blue_rings)[M].push_back(r);
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
    const size_t need = p->precision == HBX_F64 ? ring_lds_bytes<double>(maxM) : ring_lds_bytes<float>(maxM);
    if (p->ring_f64 && ring_lds_bytes<double>(maxM) > (size_t)maxlds) p->ring_f64 = false;
    if (need > (size_t)maxlds)
        return set_error(HBX_EUNSUPPORTED, "ring FFT of size " + std::to_string(maxM) +
                                               " does not fit in LDS (" + std::to_string(maxlds) +
                                               " bytes) at this precision");
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
    // threads of a ring workgroup at this plan's precision
    auto thr = [&](int M) { return p->precision == HBX_F64 ? class_threads<double>(M) : class_threads<float>(M); };
    // per-ring rotation steps (long double on the host, rounded to FP64)
    {
        std::vector<double2> rot((size_t)NROT * nr);
        const long double pi = 3.141592653589793238462643383279502884L;
        for (int r = 0; r < nr; ++r) {
            const int n = nphi[r], h = n / 2, nt = thr(ringM[r]);
            const long double ph = phi0 ? (long double)phi0[r] : 0.0L;
            auto e = [](long double a) { return make_double2((double)cosl(a), (double)sinl(a)); };
            rot[r] = e(nt * ph);
            rot[(size_t)nr + r] = e(n * ph);
            rot[2 * (size_t)nr + r] = e(h * ph);
            rot[3 * (size_t)nr + r] = e(-2.0L * pi * (long double)(nt % n) / (long double)n);
        }
        if ((rc = dalloc((void **)&p->d_rot, rot.size() * sizeof(double2)))) return rc;
        HBX_HIP(hipMemcpy(p->d_rot, rot.data(), rot.size() * sizeof(double2), hipMemcpyHostToDevice));
        if (p->ring_f64) {          // FP64 analysis transforms of an FP32 plan: FP64 thread counts
            for (int r = 0; r < nr; ++r) {
                const int n = nphi[r], nt = class_threads<double>(ringM[r]);
                const long double ph = phi0 ? (long double)phi0[r] : 0.0L;
                auto e = [](long double a) { return make_double2((double)cosl(a), (double)sinl(a)); };
                rot[r] = e(nt * ph);
                rot[3 * (size_t)nr + r] = e(-2.0L * pi * (long double)(nt % n) / (long double)n);
            }
            if ((rc = dalloc((void **)&p->d_rot64, rot.size() * sizeof(double2)))) return rc;
            HBX_HIP(hipMemcpy(p->d_rot64, rot.data(), rot.size() * sizeof(double2), hipMemcpyHostToDevice));
        }
    }
    // size classes and their twiddle tables (FP64 and FP32 roundings)
    std::map<int, int> all_M;
    for (auto &kv : pow2_rings) all_M[kv.first] = 1;
    for (auto &kv : blue_rings) all_M[kv.first] = 1;
    int64_t tw_total64 = 0, tw_total32 = 0;
    for (auto &kv : all_M) {
        FftClass c;
        c.M = kv.first;
        c.log2M = ilog2(c.M);
        c.threads = thr(c.M);
        c.tw_off64 = tw_total64;                                   // FP64 engine: 8 points per thread
        c.tw_off32 = tw_total32;                                   // FP32 engine: 16 points per thread
        tw_total64 += fft_table_entries(c.M, fft_lp<double>());
        tw_total32 += (fft_table_entries(c.M, fft_lp<float>()) + 1) & ~int64_t(1);   // rows stay 16-byte aligned
        auto up = [&](const std::vector<int32_t> &v, int32_t **d) -> int {
            if (v.empty()) return 0;
            int e = dalloc((void **)d, v.size() * sizeof(int32_t));
            if (e) return e;
            HBX_HIP(hipMemcpy(*d, v.data(), v.size() * sizeof(int32_t), hipMemcpyHostToDevice));
            return 0;
        };
        if (pow2_rings.count(c.M)) {
            c.nrings_pow2 = (int)pow2_rings[c.M].size();
            if ((rc = up(pow2_rings[c.M], &c.d_rings_pow2))) return rc;
        }
        if (blue_rings.count(c.M)) {
            c.nrings_blue = (int)blue_rings[c.M].size();
            if ((rc = up(blue_rings[c.M], &c.d_rings_blue))) return rc;
        }
        p->classes.push_back(c);
    }
    {
        std::vector<double2> t64((size_t)std::max<int64_t>(tw_total64, 1), make_double2(0.0, 0.0));
        std::vector<float2> t32((size_t)std::max<int64_t>(tw_total32, 2), make_float2(0.f, 0.f));
        for (const auto &c : p->classes) {
            fft_table_fill(c.M, fft_lp<double>(), t64.data() + c.tw_off64);
            fft_table_fill(c.M, fft_lp<float>(), t32.data() + c.tw_off32);
        }
        if ((rc = dalloc((void **)&p->d_ftw64, t64.size() * sizeof(double2)))) return rc;
        if ((rc = dalloc((void **)&p->d_ftw32, t32.size() * sizeof(float2)))) return rc;
        HBX_HIP(hipMemcpy(p->d_ftw64, t64.data(), t64.size() * sizeof(double2), hipMemcpyHostToDevice));
        HBX_HIP(hipMemcpy(p->d_ftw32, t32.data(), t32.size() * sizeof(float2), hipMemcpyHostToDevice));
    }
    auto tw_of = [&](int M, bool f64) -> int64_t {
        for (const auto &c : p->classes)
            if (c.M == M) return f64 ? c.tw_off64 : c.tw_off32;
        return 0;
    };
    if (p->precision == HBX_F64) {
        if ((rc = dalloc((void **)&p->d_V64, vtotal * sizeof(double2)))) return rc;
        for (auto &kv : hs_by_M)
            if ((rc = build_chirp_spectra<double>(kv.second, ofs_by_M[kv.first], kv.first, p->d_V64,
                                                  p->d_ftw64 + tw_of(kv.first, true))))
                return rc;
    } else {
        if ((rc = dalloc((void **)&p->d_V32, vtotal * sizeof(float2)))) return rc;
        // spectra and chirps in FP64, rounded once, where the FP64 FFT fits in LDS
        double2 *tmp = nullptr;
        if (vtotal > 0 && hipMalloc(&tmp, vtotal * sizeof(double2)) != hipSuccess)
            return set_error(HBX_ENOMEM, "hipMalloc failed for the FP64 chirp spectra");
        for (auto &kv : hs_by_M) {
            const bool f64 = ring_lds_bytes<double>(kv.first) <= (size_t)maxlds;
            if (f64)
                rc = build_chirp_spectra<double>(kv.second, ofs_by_M[kv.first], kv.first, tmp,
                                                 p->d_ftw64 + tw_of(kv.first, true));
            else
                rc = build_chirp_spectra<float>(kv.second, ofs_by_M[kv.first], kv.first, p->d_V32,
                                                p->d_ftw32 + tw_of(kv.first, false));
            if (rc) break;
            // round this class's entries (every distinct h of the class)
            for (size_t i = 0; i < kv.second.size() && f64; ++i) {
                const int64_t o = ofs_by_M[kv.first][i], len = kv.first + kv.second[i];
                hipLaunchKernelGGL(round_to_f32_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, 0,
                                   tmp + o, p->d_V32 + o, len);
            }
        }
        hipError_t e = hipDeviceSynchronize();
        if (tmp && p->ring_f64 && rc == 0) {        // kept: the FP64 analysis transforms use them
            p->d_V64 = tmp;
            p->allocations.push_back(tmp);
            p->device_bytes += vtotal * sizeof(double2);
        } else if (tmp) {
            (void)hipFree(tmp);
        }
        if (rc) return rc;
        if (e != hipSuccess) return hip_error(e, "FP32 chirp spectra");
    }
    if ((rc = plan_build_fused(p))) return rc;
    HBX_HIP(hipDeviceSynchronize());
    return 0;
}

template <typename T, typename TIn, typename TP = T>
static int ring_analysis_t(const Plan *p, int nmaps, const TIn *maps, int64_t map_stride, c2<TP> *phase,
                           const double *scale, hipStream_t s) {
    // FP64 transforms of an FP32 plan run their own launches (FP64 thread counts and tables)
    const std::vector<RingLaunch> &Ls = (sizeof(T) != sizeof(TP)) ? p->fused64 : p->fused;
    RingCall c{};
    c.maps = const_cast<TIn *>(maps);
    c.mstride = map_stride;
    c.phase = phase;
    c.scale = scale;
    c.nmaps = nmaps;
    c.unit = 8 * nmaps;
    for (const RingLaunch &L : Ls)      // every grid checked before the first launch
        if (L.size_ana * 8 * (int64_t)nmaps > INT32_MAX)
            return set_error(HBX_EUNSUPPORTED, "too many maps for one ring launch");
    timer_mark(HBX_TIMER_RING_ANALYSIS, true, s);
    for (const RingLaunch &L : Ls) {
        const int64_t grid = L.size_ana * 8 * (int64_t)nmaps;
        if (L.blue)
            hipLaunchKernelGGL((fused_ring_analysis_kernel<T, TIn, true, TP>), dim3((unsigned)grid), dim3(L.threads),
                               L.lds, s, L.d_ana, L.nseg, c);
        else
            hipLaunchKernelGGL((fused_ring_analysis_kernel<T, TIn, false, TP>), dim3((unsigned)grid),
                               dim3(L.threads), L.lds, s, L.d_ana, L.nseg, c);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_error(e, "fused_ring_analysis_kernel launch");
    }
    timer_mark(HBX_TIMER_RING_ANALYSIS, false, s);
    return 0;
}

template <typename T, typename TOut>
static int ring_synthesis_t(const Plan *p, int nmaps, const c2<T> *phase, TOut *maps, int64_t map_stride,
                            const double *scale, hipStream_t s) {
    RingCall c{};
    c.maps = maps;
    c.mstride = map_stride;
    c.phase = const_cast<c2<T> *>(phase);
    c.scale = scale;
    c.nmaps = nmaps;
    c.unit = 64 * nmaps;
    for (const RingLaunch &L : p->fused)
        if (L.size_syn * 64 * (int64_t)nmaps > INT32_MAX)
            return set_error(HBX_EUNSUPPORTED, "too many maps for one ring launch");
    timer_mark(HBX_TIMER_RING_SYNTHESIS, true, s);
    for (const RingLaunch &L : p->fused) {
        const int64_t grid = L.size_syn * 64 * (int64_t)nmaps;
        if (L.blue)
            hipLaunchKernelGGL((fused_ring_synthesis_kernel<T, TOut, true>), dim3((unsigned)grid), dim3(L.threads),
                               L.lds, s, L.d_syn, L.nseg, c);
        else
            hipLaunchKernelGGL((fused_ring_synthesis_kernel<T, TOut, false>), dim3((unsigned)grid), dim3(L.threads),
                               L.lds, s, L.d_syn, L.nseg, c);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_error(e, "fused_ring_synthesis_kernel launch");
    }
    timer_mark(HBX_TIMER_RING_SYNTHESIS, false, s);
    return 0;
}

int launch_ring_analysis(const Plan *p, int nmaps, const void *maps, int map_dtype,
                         int64_t map_stride, void *phase, const double *scale, hipStream_t s) {
    if (p->precision == HBX_F64) {
        auto *ph = static_cast<double2 *>(phase);
        if (map_dtype == HBX_DT_F64)
            return ring_analysis_t<double, double>(p, nmaps, (const double *)maps, map_stride, ph, scale, s);
        return ring_analysis_t<double, float>(p, nmaps, (const float *)maps, map_stride, ph, scale, s);
    }
    auto *ph = static_cast<float2 *>(phase);
    if (p->ring_f64) {
        if (map_dtype == HBX_DT_F64)
            return ring_analysis_t<double, double, float>(p, nmaps, (const double *)maps, map_stride, ph, scale, s);
        return ring_analysis_t<double, float, float>(p, nmaps, (const float *)maps, map_stride, ph, scale, s);
    }
    if (map_dtype == HBX_DT_F64)
        return ring_analysis_t<float, double>(p, nmaps, (const double *)maps, map_stride, ph, scale, s);
    return ring_analysis_t<float, float>(p, nmaps, (const float *)maps, map_stride, ph, scale, s);
}

int launch_ring_synthesis(const Plan *p, int nmaps, const void *phase, void *maps, int map_dtype,
                          int64_t map_stride, const double *scale, hipStream_t s) {
    if (p->precision == HBX_F64) {
        auto *ph = static_cast<const double2 *>(phase);
        if (map_dtype == HBX_DT_F64)
            return ring_synthesis_t<double, double>(p, nmaps, ph, (double *)maps, map_stride, scale, s);
        return ring_synthesis_t<double, float>(p, nmaps, ph, (float *)maps, map_stride, scale, s);
    }
    auto *ph = static_cast<const float2 *>(phase);
    if (map_dtype == HBX_DT_F64)
        return ring_synthesis_t<float, double>(p, nmaps, ph, (double *)maps, map_stride, scale, s);
    return ring_synthesis_t<float, float>(p, nmaps, ph, (float *)maps, map_stride, scale, s);
}

// ---------------------------------------------------------------------------
// mixed-geometry batches: the groups' classes merged by workgroup size, and
// (FP32 batches) the band DFT over every group's rings
// ---------------------------------------------------------------------------
int build_ring_launches(const std::vector<Plan> &plans, const std::vector<int> &nmaps,
                        const std::vector<uint64_t> &phase_off, std::vector<RingLaunch> &out, RingLaunch &band,
                        std::vector<void *> &allocs) {
    band = RingLaunch{};
    auto alloc = [&allocs](size_t bytes) -> void * {
        void *ptr = nullptr;
        if (hipMalloc(&ptr, bytes) != hipSuccess) return nullptr;
        allocs.push_back(ptr);
        return ptr;
    };
    const bool f64 = plans[0].precision == HBX_F64;
    std::map<std::pair<int, int>, std::vector<RingSeg>> by;
    for (size_t g = 0; g < plans.size(); ++g) {
        const Plan &p = plans[g];
        if (p.ring_f64) return set_error(HBX_EPLAN, "batches take plans whose ring stage runs at their precision");
        for (const FftClass &c : p.classes)
            for (int blue = 0; blue < 2; ++blue) {
                if ((blue ? c.nrings_blue : c.nrings_pow2) == 0) continue;
                RingSeg sg = class_seg(&p, c, blue, f64, p.d_rot);
                sg.nmaps = nmaps[g];
                sg.g = (int32_t)g;
                sg.phase_off = phase_off[g];
                by[{c.threads, blue}].push_back(sg);
            }
    }
    int rc = make_launches(by, f64, false, true, out, alloc);
    if (rc || f64) return rc;
    std::vector<RingSeg> segs;
    int64_t bb = 0;
    for (size_t g = 0; g < plans.size(); ++g) {
        const Plan &p = plans[g];
        RingSeg sg{};
        sg.ring_start = p.d_ring_start;
        sg.nphi = p.d_nphi;
        sg.mcut = p.d_mcut;
        sg.phi0 = p.d_phi0;
        sg.scale = p.d_scale_ana;
        sg.pstride = band_phase_elems(&p);
        sg.phase_off = phase_off[g] + band_phase_offset(&p, nmaps[g]);
        sg.ncls = p.nrings;
        sg.nmaps = nmaps[g];
        sg.g = (int32_t)g;
        sg.mmax = p.mmax;
        sg.nr = p.nrings;
        sg.block0 = (int32_t)bb;
        bb += ((int64_t)p.nrings * nmaps[g] + 7) / 8 * 8;
        if (bb > INT32_MAX) return set_error(HBX_EUNSUPPORTED, "batch too large for one band launch");
        segs.push_back(sg);
    }
    band.threads = RBV_LANES;
    band.nseg = (int)segs.size();
    band.size_ana = bb;
    void *d = alloc(segs.size() * sizeof(RingSeg));
    if (!d) return set_error(HBX_ENOMEM, "hipMalloc for the band launch table failed");
    HBX_HIP(hipMemcpy(d, segs.data(), segs.size() * sizeof(RingSeg), hipMemcpyHostToDevice));
    band.d_ana = static_cast<RingSeg *>(d);
    return HBX_OK;
}

static RingCall batch_call(const void *const *maps_tab, const int64_t *mstride_tab, char *ws) {
    RingCall c{};
    c.maps_tab = maps_tab;
    c.mstride_tab = mstride_tab;
    c.ws = ws;
    c.unit = 1;
    return c;
}

int launch_multi_ring_analysis(const RingLaunch &L, int precision, int map_dtype, const void *const *maps_tab,
                               const int64_t *mstride_tab, char *ws, hipStream_t s) {
    const RingCall c = batch_call(maps_tab, mstride_tab, ws);
#define HBX_MRA(T, TI, B)                                                                                          \
    hipLaunchKernelGGL((fused_ring_analysis_kernel<T, TI, B, T>), dim3((unsigned)L.size_ana), dim3(L.threads), L.lds, \
                       s, L.d_ana, L.nseg, c)
    const bool din = map_dtype == HBX_DT_F64;
    if (precision == HBX_F64) {
        if (L.blue) { if (din) HBX_MRA(double, double, true); else HBX_MRA(double, float, true); }
        else { if (din) HBX_MRA(double, double, false); else HBX_MRA(double, float, false); }
    } else {
        if (L.blue) { if (din) HBX_MRA(float, double, true); else HBX_MRA(float, float, true); }
        else { if (din) HBX_MRA(float, double, false); else HBX_MRA(float, float, false); }
    }
#undef HBX_MRA
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "fused_ring_analysis_kernel (batch) launch");
    return HBX_OK;
}

int launch_multi_ring_synthesis(const RingLaunch &L, int precision, int map_dtype, void *const *maps_tab,
                                const int64_t *mstride_tab, char *ws, hipStream_t s) {
    const RingCall c = batch_call(const_cast<const void *const *>(maps_tab), mstride_tab, ws);
#define HBX_MRS(T, TO, B)                                                                                       \
    hipLaunchKernelGGL((fused_ring_synthesis_kernel<T, TO, B>), dim3((unsigned)L.size_syn), dim3(L.threads), L.lds, \
                       s, L.d_syn, L.nseg, c)
    const bool dout = map_dtype == HBX_DT_F64;
    if (precision == HBX_F64) {
        if (L.blue) { if (dout) HBX_MRS(double, double, true); else HBX_MRS(double, float, true); }
        else { if (dout) HBX_MRS(double, double, false); else HBX_MRS(double, float, false); }
    } else {
        if (L.blue) { if (dout) HBX_MRS(float, double, true); else HBX_MRS(float, float, true); }
        else { if (dout) HBX_MRS(float, double, false); else HBX_MRS(float, float, false); }
    }
#undef HBX_MRS
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "fused_ring_synthesis_kernel (batch) launch");
    return HBX_OK;
}

int launch_multi_ring_band(const RingLaunch &L, int map_dtype, const void *const *maps_tab,
                           const int64_t *mstride_tab, char *ws, hipStream_t s) {
    if (L.nseg == 0) return HBX_OK;
    const RingCall c = batch_call(maps_tab, mstride_tab, ws);
    if (map_dtype == HBX_DT_F64)
        hipLaunchKernelGGL(multi_ring_band_kernel<double>, dim3((unsigned)L.size_ana), dim3(RBV_LANES), 0, s,
                           L.d_ana, L.nseg, c);
    else
        hipLaunchKernelGGL(multi_ring_band_kernel<float>, dim3((unsigned)L.size_ana), dim3(RBV_LANES), 0, s,
                           L.d_ana, L.nseg, c);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "multi_ring_band_kernel launch");
    return HBX_OK;
}

}  // namespace hbx
