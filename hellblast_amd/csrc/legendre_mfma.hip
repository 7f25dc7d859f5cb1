// Legendre synthesis on the matrix cores (FP32 plans, batches of maps).
//
// Per m the synthesis is a product  G[ring pair][n] = sum_l lambda_l(x_p) at[l][n]
// with n = (map, re/im) over a group of 16 maps and at = s_lm * a_lm.  It maps
// onto v_mfma_f32_32x32x2f32 (D[32x32] += A[32x2] B[2x32]):
//   A[row = ring][k] = lambda_{l0+2k}(x_ring)   (FP64 recursion, cvt to FP32)
//   B[k][col = n]    = at_{l0+2k}[n]            (staged in LDS per 64-l chunk)
// one MFMA for the even l of a 4-l step and one for the odd l, so the two
// accumulators are G_even and G_odd and the mirror rings follow as
// G_n = G_even + G_odd, G_s = G_even - G_odd.
// A wave covers two 32-ring tiles: lane t runs the recursion of ring t of the
// wave (tile t/32) for l0..l0+3, and v_permlane32_swap exchanges the halves
// so each tile's A operand gets l0 (lanes 0-31) and l0+2 (lanes 32-63):
//   swap(v0, v2) -> (tile 0: [v0 | v2 of lanes 0-31], tile 1: [v0 of lanes
//   32-63 | v2]), and likewise (v1, v3) for the odd l.
// Every ring's recursion runs once, and the staged B operands serve 64 rings.
// Lane layout of the 32x32x2 f32 instruction (checked on gfx950):
//   A: lane = row + 32 k     B: lane = col + 32 k
//   D[row][col]: lane = col + 32 ((row/4)%2), reg = (row%4) + 4 (row/8)
// and of v_permlane32_swap(a, b) (checked on gfx950): returns
//   ([a lanes 0-31 | b lanes 0-31], [a lanes 32-63 | b lanes 32-63]).
// Reference counterpart: the per-(ring, m) l-sum of isht.py:15-21.
#include "hbx_device.h"
#include "hbx_internal.h"

namespace hbx {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int SYN_LCHUNK = 64;     // l values staged per chunk
constexpr int SYN_WAVES = 4;       // waves per workgroup (same m)
constexpr int SYN_PAIRS_WG = SYN_WAVES * 64;
constexpr int NEVER_L = 0x7fffffff;

__device__ __forceinline__ int wmin64(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int wmax64(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return __builtin_amdgcn_readfirstlane(v);
}

// (lo, hi) halves exchange: returns {[a.lo | b.lo], [a.hi | b.hi]}
__device__ __forceinline__ void swap_halves(float a, float b, float *t0, float *t1) {
    auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    *t0 = __uint_as_float(r[0]);
    *t1 = __uint_as_float(r[1]);
}

// (the waves-per-EU bound lets the four 32x32 accumulators live in VGPRs;
// without it the compiler cycles them through one AGPR block with copies)
__global__ void __launch_bounds__(SYN_WAVES * 64) __attribute__((amdgpu_waves_per_eu(2, 4)))
leg_synthesis_mfma_kernel(const float2 *__restrict__ alm, int64_t astride, float *__restrict__ phase,
                          const double *__restrict__ px, const int64_t *__restrict__ rofs,
                          const double *__restrict__ alpha, const double *__restrict__ sl,
                          const int32_t *__restrict__ lseed, const double2 *__restrict__ seed, int npairs, int nr,
                          int lmax, int mmax, int nmb, int nsb, int nmaps, int adjoint) {
    // per-chunk operands: at[l][n] and alpha_l
    __shared__ float stage[SYN_LCHUNK][33];
    __shared__ double alds[SYN_LCHUNK];
    __shared__ int wg_start;
    int mb, item;
    xcd_map(blockIdx.x, MB * nsb, &mb, &item);
    const int m = mb * MB + item / nsb;
    const int sb = item % nsb;
    if (m > mmax) return;
    const int g = blockIdx.y;                       // group of 16 maps
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int r = lane & 31, h = lane >> 5;

    // this lane's ring pair: ring lane of the wave's 64
    const int pbase = (sb * SYN_WAVES + w) * 64;
    const int p = pbase + lane;
    double x = 0.0, u1 = 0.0, u2 = 0.0;
    double2 sd = make_double2(0.0, 0.0);
    int ls = NEVER_L;
    if (p < npairs) {
        x = px[p];
        ls = lseed[(int64_t)m * npairs + p];
        if (ls <= lmax) sd = seed[(int64_t)m * npairs + p];
    }
    const int wstart = wmin64(ls);
    const int winj = wmax64(ls <= lmax ? ls : -1);
    if (tid == 0) wg_start = NEVER_L;
    __syncthreads();
    if (lane == 0) atomicMin(&wg_start, wstart);
    __syncthreads();
    const int lstart = wg_start;

    const int64_t ro = rofs[m];
    const double *al = alpha + ro - m;      // al[l] = alpha_lm (zero-padded past lmax)
    const double *sv = sl + ro - m;
    const double fadj = (adjoint && m > 0) ? 0.5 : 1.0;   // map2alm transpose: g/c_m
    const int64_t i0 = alm_index(m, m, lmax) - m;   // alm[b][i0 + l]
    // stage slots of this thread: l = c0 + (tid & 63), maps b = tid/64 + 4 q
    constexpr int PER = 16 / SYN_WAVES;
    const int sll = tid & 63, sb0 = tid >> 6;
    const float2 *abase = alm + i0 + sll;
    auto fetch = [&](int c0, float2 *v, double *a) {
        const int l = c0 + sll;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int b = g * 16 + sb0 + SYN_WAVES * q;
            float2 val = make_float2(0.f, 0.f);
            if (l <= lmax && b < nmaps) {
                const float2 av = abase[(int64_t)b * astride + c0];
                const double f = sv[l] * fadj;
                val = make_float2((float)(av.x * f), (float)(av.y * f));
            }
            v[q] = val;
        }
        *a = (tid < SYN_LCHUNK) ? al[c0 + tid] : 0.0;
    };
    auto put = [&](const float2 *v, double a) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int n = 2 * (sb0 + SYN_WAVES * q);
            stage[sll][n] = v[q].x;
            stage[sll][n + 1] = v[q].y;
        }
        if (tid < SYN_LCHUNK) alds[tid] = a;
    };

    floatx16 De[2], Do[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        De[j] = floatx16{0};
        Do[j] = floatx16{0};
    }
    // one buffer: the chunk's operands arrive in a single round trip and
    // co-resident workgroups hide it
    for (int c0 = lstart; c0 <= lmax; c0 += SYN_LCHUNK) {
        {
            float2 pv[PER];
            double pa;
            fetch(c0, pv, &pa);
            put(pv, pa);
        }
        __syncthreads();
        if (c0 + SYN_LCHUNK > wstart) {
            for (int kk = 0; kk < SYN_LCHUNK / 4; ++kk) {
                const int l0 = c0 + 4 * kk;
                if (l0 > lmax) break;
                const double a0 = alds[4 * kk], a1 = alds[4 * kk + 1];
                const double a2 = alds[4 * kk + 2], a3 = alds[4 * kk + 3];
                if (l0 <= winj && ls == l0) {
                    u2 = sd.x;
                    u1 = sd.y;
                }
                const double v0 = fma(a0 * x, u1, -u2);
                const double v1 = fma(a1 * x, v0, -u1);
                const double v2 = fma(a2 * x, v1, -v0);
                const double v3 = fma(a3 * x, v2, -v1);
                u2 = v2;
                u1 = v3;
                float ae0, ae1, ao0, ao1;
                swap_halves((float)v0, (float)v2, &ae0, &ae1);
                swap_halves((float)v1, (float)v3, &ao0, &ao1);
                const float be = stage[4 * kk + 2 * h][r];
                const float bo = stage[4 * kk + 2 * h + 1][r];
                De[0] = __builtin_amdgcn_mfma_f32_32x32x2f32(ae0, be, De[0], 0, 0, 0);
                Do[0] = __builtin_amdgcn_mfma_f32_32x32x2f32(ao0, bo, Do[0], 0, 0, 0);
                De[1] = __builtin_amdgcn_mfma_f32_32x32x2f32(ae1, be, De[1], 0, 0, 0);
                Do[1] = __builtin_amdgcn_mfma_f32_32x32x2f32(ao1, bo, Do[1], 0, 0, 0);
            }
        }
        __syncthreads();
    }

    // D[row = ring][col = n]: this lane holds col = r, rows (q%4) + 8 (q/4) + 4 h.
    // Ring pair p couples ring p (north) and ring nr-1-p (south).
    const int b = g * 16 + (r >> 1), c = r & 1;
    if (b >= nmaps) return;
    float *out = phase + (((int64_t)b * nmb + mb) * nr) * (2 * MB) + (m & 7) * 2 + c;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int base = pbase + 32 * j;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int row = (q & 3) + 8 * (q >> 2) + 4 * h;
            const int pp = base + row;
            if (pp >= npairs) continue;
            const float ge = De[j][q], go = Do[j][q];
            const int rn = pp, rs = nr - 1 - pp;
            out[(int64_t)rn * (2 * MB)] = ge + go;
            if (rs != rn) out[(int64_t)rs * (2 * MB)] = ge - go;
        }
    }
}

int launch_leg_synthesis_mfma(const Plan *p, int nmaps, const void *alm, int64_t astride, void *phase,
                              int adjoint, hipStream_t s) {
    const int nsb = (p->npairs + SYN_PAIRS_WG - 1) / SYN_PAIRS_WG;
    const int nmb8 = (p->nmb + 7) / 8 * 8;
    dim3 grid(nmb8 * MB * nsb, (nmaps + 15) / 16);
    timer_mark(HBX_TIMER_LEGENDRE_SYNTHESIS, true, s);
    hipLaunchKernelGGL(leg_synthesis_mfma_kernel, grid, dim3(SYN_WAVES * 64), 0, s, (const float2 *)alm, astride,
                       (float *)phase, p->d_px, p->d_rofs, p->d_alpha, p->d_sl,
                       p->d_lseed, p->d_seed, p->npairs, p->nrings, p->lmax, p->mmax, p->nmb, nsb, nmaps, adjoint);
    timer_mark(HBX_TIMER_LEGENDRE_SYNTHESIS, false, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "leg_synthesis_mfma_kernel launch");
    return 0;
}

// ---------------------------------------------------------------------------
// Legendre analysis on the matrix cores (FP32 plans, batches of maps).
//
// Per m:  a[n][l] = sum_p S_par[p][n] lambda_l(x_p), n = (map, re/im) over 16
// maps, par = parity of l-m (S_even = F_n + F_s, S_odd = F_n - F_s).  With
// v_mfma_f32_16x16x4f32 (D[16x16] += A[16x4] B[4x16]) the ring pairs are the
// contraction index:
//   A[row = n][k] = S_par[ring k][n]             (held in registers all kernel)
//   B[k][col]     = lambda_{c0+2col+par}(ring k) (FP64 recursion -> FP32 ->
//                                                 wave-private LDS transpose)
// A workgroup is W waves x 64 ring pairs of one m (W = 4, 8 or 16, a plan
// tile).  After its MFMAs each wave parks its D tiles in its own (consumed)
// lambda tile; the threads then sum the chunk's 32 l x 32 n outputs over the
// W waves in a fixed order and write them.  (LDS float atomics for this sum
// measured 4x slower.)
// Lane layout of the 16x16x4 f32 instruction (checked on gfx950):
//   A: lane = row + 16 k   B: lane = col + 16 k   D: lane = col + 16 (row/4), reg = row%4
// Reference counterpart: the per-(m, l) ring sum of sht_raw.py:53 /
// sht_fallback.py:63.
// ---------------------------------------------------------------------------
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int ANA_L = 32;                 // l per chunk: 16 even + 16 odd columns
// A wave's lambda tile holds, per parity and column (16 l of that parity),
// the 64 ring pairs q at (q%4)*16 + q/4: one recursion step writes 64
// consecutive floats, and the lane that feeds B[k = q%4][col] reads its 16
// values (q/4 = s = 0..15) as four 16-byte loads.  Rows are padded to 68
// floats so the 16 lanes of a quarter wave hit disjoint banks.
constexpr int ANA_COL = 68;
constexpr int ANA_PAR = 16 * ANA_COL;
constexpr int ANA_LAM = 2 * ANA_PAR;      // floats of one wave's lambda tile
constexpr size_t ana_lds(int waves) { return (size_t)(waves * ANA_LAM + 4 * ANA_L + 4) * sizeof(float); }

template <int ANA_WAVES>
__global__ void __launch_bounds__(ANA_WAVES * 64)
leg_analysis_mfma_kernel(const float2 *__restrict__ phase, int64_t pstride, float2 *__restrict__ alm,
                         int64_t astride, float2 *__restrict__ partial, const double *__restrict__ px,
                         const int64_t *__restrict__ rofs, const double *__restrict__ alpha,
                         const double *__restrict__ sl, const int32_t *__restrict__ lseed,
                         const double2 *__restrict__ seed, int npairs, int nr, int lmax, int mmax, int ngroups,
                         int64_t nlm, int nmaps, int adjoint, int mbeg) {
    extern __shared__ __align__(16) float smem[];
    int mb, item;
    xcd_map(blockIdx.x, MB * ngroups, &mb, &item);
    const int m = mb * MB + item / ngroups;
    const int grp = item % ngroups;
    if (m > mmax || m < mbeg) return;
    const int g = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    float *lam = smem + w * ANA_LAM;
    double *alds = reinterpret_cast<double *>(smem + ANA_WAVES * ANA_LAM);   // [2][32]
    int &wg_start = *reinterpret_cast<int *>(smem + ANA_WAVES * ANA_LAM + 4 * ANA_L);

    const int pw = grp * ANA_WAVES * 64 + w * 64;       // first ring pair of this wave
    const int p = pw + lane;
    double x = 0.0, u1 = 0.0, u2 = 0.0;
    double2 sd = make_double2(0.0, 0.0);
    int ls = NEVER_L;
    if (p < npairs) {
        x = px[p];
        ls = lseed[(int64_t)m * npairs + p];
        if (ls <= lmax) sd = seed[(int64_t)m * npairs + p];
    }
    const int wstart = wmin64(ls);
    const int winj = wmax64(ls <= lmax ? ls : -1);

    // A operands: S_par[ring pw + 4 s + lane/16][n = 16 tile + lane%16];
    // ring pair q couples ring q (north) and ring nr-1-q (south)
    float Sv[16][2][2];
    {
        const int c = lane & 1;
        const float2 *ph = phase + (int64_t)mb * nr * MB + (m & 7);
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            const int pq = pw + 4 * s + (lane >> 4);
            const int pp = min(pq, npairs - 1);
            const bool self = (nr - 1 - pp) == pp;
#pragma unroll
            for (int tile = 0; tile < 2; ++tile) {
                const int bq = g * 16 + tile * 8 + ((lane & 15) >> 1);
                const float2 *pb = ph + (int64_t)min(bq, nmaps - 1) * pstride;
                const float2 a = pb[(int64_t)pp * MB];
                const float2 z = self ? make_float2(0.f, 0.f) : pb[(int64_t)(nr - 1 - pp) * MB];
                const bool ok = pq < npairs && bq < nmaps;
                Sv[s][tile][0] = ok ? (c ? a.y + z.y : a.x + z.x) : 0.0f;
                Sv[s][tile][1] = ok ? (c ? a.y - z.y : a.x - z.x) : 0.0f;
            }
        }
    }
    if (tid == 0) wg_start = NEVER_L;
    __syncthreads();
    if (lane == 0) atomicMin(&wg_start, wstart);
    __syncthreads();
    const int lstart = wg_start;

    const bool direct = (ngroups == 1);
    const int64_t i0 = alm_index(m, m, lmax) - m;       // + l
    const int64_t ro = rofs[m];
    const double *al = alpha + ro - m;                  // zero-padded past lmax
    const double *sv = sl + ro - m;
    auto store = [&](int l, int n, float v) {
        const int b = g * 16 + (n >> 1), c = n & 1;
        if (b >= nmaps) return;
        if (direct) {
            double d = (double)v * sv[l];
            if (adjoint) {
                if (m > 0) d *= 2.0;
                else if (c == 1) d = 0.0;
            }
            reinterpret_cast<float *>(alm + (int64_t)b * astride + i0 + l)[c] = (float)d;
        } else {
            reinterpret_cast<float *>(partial + ((int64_t)b * ngroups + grp) * nlm + i0 + l)[c] = v;
        }
    };
    // coefficients below every ring pair's switch-on are zero
    const int lz = lstart > lmax ? lmax + 1 : lstart;
    for (int t = tid; t < (lz - m) * 32; t += ANA_WAVES * 64) store(m + (t >> 5), t & 31, 0.0f);
    if (lstart > lmax) return;


    if (tid < ANA_L) alds[tid] = al[lstart + tid];
    __syncthreads();
    int buf = 0;
    for (int c0 = lstart; c0 <= lmax; c0 += ANA_L, buf ^= 1) {
        const double *ab = alds + buf * ANA_L;
        const double anext = (tid < ANA_L) ? al[c0 + ANA_L + tid] : 0.0;   // in flight this chunk
        floatx4 D[2][2];
#pragma unroll
        for (int tile = 0; tile < 2; ++tile)
#pragma unroll
            for (int par = 0; par < 2; ++par) D[tile][par] = floatx4{0.f, 0.f, 0.f, 0.f};
        if (c0 + ANA_L > wstart) {
            // lambda for this lane's ring over l = c0 .. c0+31 into the wave tile
#pragma unroll
            for (int t = 0; t < ANA_L; ++t) {
                if ((t & 7) == 0 && c0 <= winj && ls == c0 + t) {
                    u2 = sd.x;
                    u1 = sd.y;
                }
                const double u = fma(ab[t] * x, u1, -u2);
                u2 = u1;
                u1 = u;
                lam[(t & 1) * ANA_PAR + (t >> 1) * ANA_COL + (lane & 3) * 16 + (lane >> 2)] = (float)u;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int par = 0; par < 2; ++par) {
                // B[k = lane/16][col = lane%16] for s = 0..15: ring pair 4 s + k
                const float4 *bp =
                    reinterpret_cast<const float4 *>(lam + par * ANA_PAR + (lane & 15) * ANA_COL + (lane >> 4) * 16);
                float bv[16];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float4 q4 = bp[j];
                    bv[4 * j] = q4.x;
                    bv[4 * j + 1] = q4.y;
                    bv[4 * j + 2] = q4.z;
                    bv[4 * j + 3] = q4.w;
                }
#pragma unroll
                for (int s = 0; s < 16; ++s)
#pragma unroll
                    for (int tile = 0; tile < 2; ++tile)
                        D[tile][par] = __builtin_amdgcn_mfma_f32_16x16x4f32(Sv[s][tile][par], bv[s], D[tile][par], 0, 0, 0);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        // park D in this wave's (now consumed) lambda tile:
        // [(tile*2 + par)*4 + reg][lane], n = 16 tile + 4 (lane/16) + reg, l = c0 + 2 (lane%16) + par
#pragma unroll
        for (int tile = 0; tile < 2; ++tile)
#pragma unroll
            for (int par = 0; par < 2; ++par)
#pragma unroll
                for (int reg = 0; reg < 4; ++reg) lam[((tile * 2 + par) * 4 + reg) * 64 + lane] = D[tile][par][reg];
        if (tid < ANA_L) alds[(buf ^ 1) * ANA_L + tid] = anext;
        __syncthreads();
        // output o = (n = o/32, l = c0 + o%32), parked by each wave at ridx
#pragma unroll
        for (int o = tid; o < 32 * ANA_L; o += ANA_WAVES * 64) {
            const int on = o >> 5, olo = o & 31;
            const int ridx = (((on >> 4) * 2 + (olo & 1)) * 4 + (on & 3)) * 64 + (olo >> 1) + 16 * ((on & 15) >> 2);
            float v = 0.0f;
#pragma unroll
            for (int q = 0; q < ANA_WAVES; ++q) v += smem[q * ANA_LAM + ridx];
            const int l = c0 + olo;
            if (l <= lmax) store(l, on, v);
        }
        __syncthreads();
    }
}

// waves per workgroup: the plan's tile, else 8 (4 for small geometries, whose
// ring pairs would leave most of a larger workgroup idle)
int mfma_analysis_waves(const Plan *p) {
    if (p->tile_mfma_waves > 0) return p->tile_mfma_waves;
    return p->npairs <= 256 ? 4 : 8;
}

int mfma_analysis_groups(const Plan *p) {
    const int pairs_wg = mfma_analysis_waves(p) * 64;
    return (p->npairs + pairs_wg - 1) / pairs_wg;
}

int launch_leg_analysis_mfma(const Plan *p, int nmaps, const void *phase, void *alm, int64_t astride, void *partial,
                             int adjoint, hipStream_t s) {
    const int mbeg = std::min(M_LO, p->mmax + 1);    // m < M_LO run in the FP64 band
    if (mbeg > p->mmax) return 0;
    const int W = mfma_analysis_waves(p);
    const int G = mfma_analysis_groups(p);
    const int nmb8 = (p->nmb + 7) / 8 * 8;
    dim3 grid(nmb8 * MB * G, (nmaps + 15) / 16);
    const int64_t pstride = phase_elems_per_map(p);
    timer_mark(HBX_TIMER_LEGENDRE_ANALYSIS, true, s);
#define HBX_AM(WV)                                                                                            \
    hipLaunchKernelGGL(leg_analysis_mfma_kernel<WV>, grid, dim3(WV * 64), ana_lds(WV), s, (const float2 *)phase, \
                       pstride, (float2 *)alm, astride, (float2 *)partial, p->d_px, p->d_rofs, p->d_alpha,     \
                       p->d_sl, p->d_lseed, p->d_seed, p->npairs, p->nrings, p->lmax, p->mmax, G, p->nlm, nmaps, \
                       adjoint, mbeg)
    switch (W) {
    case 4: HBX_AM(4); break;
    case 8: HBX_AM(8); break;
    case 16: HBX_AM(16); break;
    default: return set_error(HBX_EUNSUPPORTED, "matrix-core analysis waves must be 4, 8 or 16");
    }
#undef HBX_AM
    timer_mark(HBX_TIMER_LEGENDRE_ANALYSIS, false, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "leg_analysis_mfma_kernel launch");
    if (G > 1) return launch_group_reduce(p, nmaps, partial, alm, astride, G, p->nlm, mbeg, p->mmax, false, adjoint, s);
    return 0;
}

// raise the dynamic-LDS limit of the analysis kernels (called at plan creation)
int prepare_mfma_kernels() {
    const void *k[3] = {reinterpret_cast<const void *>(leg_analysis_mfma_kernel<4>),
                        reinterpret_cast<const void *>(leg_analysis_mfma_kernel<8>),
                        reinterpret_cast<const void *>(leg_analysis_mfma_kernel<16>)};
    const int wv[3] = {4, 8, 16};
    for (int i = 0; i < 3; ++i) {
        hipError_t e = hipFuncSetAttribute(k[i], hipFuncAttributeMaxDynamicSharedMemorySize, (int)ana_lds(wv[i]));
        if (e != hipSuccess) return hip_error(e, "hipFuncSetAttribute(leg_analysis_mfma_kernel)");
    }
    return 0;
}

}  // namespace hbx
