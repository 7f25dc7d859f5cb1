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

// fp16 pairs for the split-fp16 matrix-core kernels: v = hi + lo,
// hi = fp16(v), lo = fp16(v - hi)
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pack_h2(half2v h) { return __builtin_bit_cast(uint32_t, h); }

// lo = fp16(v - hi), the residual as one mixed-precision FMA per value
// (v_fma_mix_f32 reads the fp16 halves directly); neg1 = -1 arrives at run
// time so the compiler keeps the FMA form
__device__ __forceinline__ half2v split_lo(float2v v, half2v hi, float neg1) {
    const float r0 = __builtin_fmaf((float)hi[0], neg1, v[0]);
    const float r1 = __builtin_fmaf((float)hi[1], neg1, v[1]);
    return __builtin_convertvector((float2v){r0, r1}, half2v);
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
    // synthesis layout (floats): [map][m/8][half][pair/8][m%8][pair%8][re, im]
    const int64_t sbe = syn_block_elems(npairs), south = 2 * (int64_t)pairs8(npairs) * MB;
    float *out = phase + ((int64_t)b * nmb + mb) * sbe * 2 + (m & 7) * 16 + c;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int base = pbase + 32 * j;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int row = (q & 3) + 8 * (q >> 2) + 4 * h;
            const int pp = base + row;
            if (pp >= npairs) continue;
            const float ge = De[j][q], go = Do[j][q];
            const int64_t o = 2 * ((pp >> 3) * (8 * MB) + (pp & 7));
            out[o] = ge + go;
            if (nr - 1 - pp != pp) out[south + o] = ge - go;
        }
    }
}

// ---------------------------------------------------------------------------
// Split-fp16 matrix-core synthesis: the product of leg_synthesis_mfma_kernel
// on v_mfma_f32_32x32x16_f16 with both operands as fp16 hi/lo pairs (three
// products, FP32 accumulation; see leg_analysis_h16_kernel):
//   A[row = ring][k]  = mu_l(x_ring), l = l0 + 2 k + par (16 same-parity l
//                       of a 32-l step; lane = row + 32 (k/8), halves k%8)
//   B[k][col = n]     = 2^k_b s_lm a_lm  (LDS, per 64-l chunk)
// Each lane runs the FP64 recursion of its ring pair over the 32 l of a step;
// the first 16 l form block 0, the next 16 block 1, and v_permlane32_swap of
// (block 0, block 1) yields the A operands of both 32-ring tiles, so every
// recursion runs once.  Range: per (map, m) the alm row is scaled by 2^k_b
// so its maximum lies in [2^13, 2^14) (syn_exp_kernel), and the output
// columns are unscaled by 2^-k_b (exact).
// Lane layout of the 32x32x16 f16 instruction (checked on gfx950):
//   A: lane = row + 32 (k/8), halves k%8   B: lane = col + 32 (k/8), halves k%8
//   D[row][col]: lane = col + 32 ((row/4)%2), reg = (row%4) + 4 (row/8)
// Reference counterpart: the per-(ring, m) l-sum of isht.py:15-21.
// ---------------------------------------------------------------------------
typedef _Float16 half8s __attribute__((ext_vector_type(8)));

constexpr int SH_WAVES = 4;                     // waves per workgroup (same m), 64 ring pairs each
constexpr int SH_LCHUNK = 64;                   // l staged per chunk (two 32-l steps)
constexpr int SH_BCOL = 40;                     // halves per staged column: 32 + an 8-half skew

// Synthesis operands, prepared once per call: per (m, group of 16 maps) the
// scaled alm rows 2^k_b s_lm a_lm fadj (k_b: the row maximum lands in
// [2^13, 2^14)) as fp16 hi/lo pairs, in 64-l chunks laid out as the kernel
// stages them, [hi/lo][parity][n = 2 map + re/im][same-parity l] (8 KB per
// chunk); chunk j of row m starts at l = m + 64 j, at chunk index cofs[m] + j.
constexpr int SH_CHUNK_HALVES = 2 * 2 * 32 * 32;

__global__ void __launch_bounds__(256)
syn_prep_kernel(const float2 *__restrict__ alm, int64_t astride, const int64_t *__restrict__ rofs,
                const double *__restrict__ sl, const int64_t *__restrict__ cofs, int64_t nchunks, int lmax,
                int mmax, int nmaps, int adjoint, float neg1, _Float16 *__restrict__ prep, int *__restrict__ kexp) {
    __shared__ double red[4][16];
    __shared__ double scs[16];
    __shared__ __align__(16) _Float16 cbuf[SH_CHUNK_HALVES];
    const int m = blockIdx.x, g = blockIdx.y, tid = threadIdx.x, lane = tid & 63;
    const int64_t i0 = alm_index(m, m, lmax) - m;
    const double *sv = sl + rofs[m] - m;
    const double fadj = (adjoint && m > 0) ? 0.5 : 1.0;
    const int nb = min(16, nmaps - g * 16);          // maps of this group
    const float2 *arow = alm + (int64_t)(g * 16) * astride + i0;
    // 1. row maxima of the 16 maps (coalesced along l), in FP64 so that every
    // finite FP32 row keeps its full precision
    {
        double mx[16];
#pragma unroll
        for (int bi = 0; bi < 16; ++bi) mx[bi] = 0.0;
        for (int l = m + tid; l <= lmax; l += blockDim.x) {
            const double f = fabs(sv[l] * fadj);
#pragma unroll
            for (int bi = 0; bi < 16; ++bi)
                if (bi < nb) {
                    const float2 a = arow[(int64_t)bi * astride + l];
                    mx[bi] = fmax(mx[bi], (double)fmaxf(fabsf(a.x), fabsf(a.y)) * f);
                }
        }
#pragma unroll
        for (int bi = 0; bi < 16; ++bi) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) mx[bi] = fmax(mx[bi], __shfl_xor(mx[bi], o, 64));
            if (lane == 0) red[tid >> 6][bi] = mx[bi];
        }
    }
    __syncthreads();
    if (tid < 16) {
        // row maximum scaled to [2^13, 2^14); the scale rides on the FP64
        // factor s_lm, so no FP32 intermediate can leave the FP32 range
        const double mx = fmax(fmax(red[0][tid], red[1][tid]), fmax(red[2][tid], red[3][tid]));
        int e = 0;
        (void)frexp(mx, &e);
        const int k = mx > 0.0 ? 14 - e : 0;
        scs[tid] = tid < nb ? ldexp(fadj, k) : 0.0;
        if (tid < nb) kexp[(int64_t)(g * 16 + tid) * (mmax + 1) + m] = k;
    }
    __syncthreads();
    // 2. per 64-l chunk: element e = tid + 256 q -> (map e / 64, l = c0 + e % 64), split into
    // the LDS chunk [hl][par][n][kk], then 16-byte pieces to the prepared buffer
    const int nch = (lmax - m) / SH_LCHUNK + 1;
    _Float16 *rowp = prep + ((int64_t)g * nchunks + cofs[m]) * SH_CHUNK_HALVES;
    for (int j = 0; j < nch; ++j) {
        const int c0 = m + SH_LCHUNK * j;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + 256 * q, bi = e >> 6, ll = e & 63, l = c0 + ll;
            float2v v = {0.0f, 0.0f};
            if (l <= lmax && bi < nb) {
                const float2 a = arow[(int64_t)bi * astride + l];
                const double f = sv[l] * scs[bi];
                v[0] = (float)(a.x * f);
                v[1] = (float)(a.y * f);
            }
            const half2v h = __builtin_convertvector(v, half2v);
            const half2v lo = split_lo(v, h, neg1);
            const int par = ll & 1, kk = ll >> 1;
            cbuf[((0 * 2 + par) * 32 + 2 * bi) * 32 + kk] = h[0];
            cbuf[((0 * 2 + par) * 32 + 2 * bi + 1) * 32 + kk] = h[1];
            cbuf[((1 * 2 + par) * 32 + 2 * bi) * 32 + kk] = lo[0];
            cbuf[((1 * 2 + par) * 32 + 2 * bi + 1) * 32 + kk] = lo[1];
        }
        __syncthreads();
        _Float16 *cp = rowp + (int64_t)j * SH_CHUNK_HALVES;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int pc = tid + 256 * q;
            *reinterpret_cast<uint4 *>(cp + 8 * pc) = *reinterpret_cast<const uint4 *>(cbuf + 8 * pc);
        }
        __syncthreads();
    }
}

// per-m chunk offsets of the prepared operands (plan time)
int plan_build_syn_chunks(Plan *p) {
    std::vector<int64_t> cofs(p->mmax + 2, 0);
    for (int m = 0; m <= p->mmax; ++m) cofs[m + 1] = cofs[m] + ((p->lmax - m) / SH_LCHUNK + 1);
    p->syn_chunks = cofs[p->mmax + 1];
    hipError_t e = hipMalloc(&p->d_cofs, cofs.size() * sizeof(int64_t));
    if (e != hipSuccess) return set_error(HBX_ENOMEM, "hipMalloc failed for synthesis chunk offsets");
    p->allocations.push_back(p->d_cofs);
    p->device_bytes += cofs.size() * sizeof(int64_t);
    HBX_HIP(hipMemcpy(p->d_cofs, cofs.data(), cofs.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    return 0;
}

// workspace of the split-fp16 synthesis: prepared operands | k_b(m)
uint64_t syn_prep_bytes(const Plan *p, int nmaps) {
    const uint64_t ng = (uint64_t)(nmaps + 15) / 16;
    const uint64_t a = (ng * (uint64_t)p->syn_chunks * SH_CHUNK_HALVES * 2 + 255) & ~uint64_t(255);
    return a + (uint64_t)nmaps * (p->mmax + 1) * sizeof(int);
}

__device__ __forceinline__ uint32_t h2bits(half2v h) { return __builtin_bit_cast(uint32_t, h); }

// NG groups of 16 maps per workgroup share each recursion (NG = 2: twice the
// matrix work per lambda value; the accumulators double)
template <int NG>
__global__ void __launch_bounds__(SH_WAVES * 64) __attribute__((amdgpu_waves_per_eu(2, 4)))
leg_synthesis_h16_kernel(const float2 *__restrict__ alm, int64_t astride, float *__restrict__ phase,
                         const double *__restrict__ px, const int64_t *__restrict__ rofs,
                         const double *__restrict__ alpha, const double *__restrict__ sl,
                         const int32_t *__restrict__ lseed, const double2 *__restrict__ seed, int npairs, int nr,
                         int lmax, int mmax, int nmb, int nsb, int nmaps, const int *__restrict__ kexp,
                         const _Float16 *__restrict__ prep, const int64_t *__restrict__ cofs, int64_t nchunks,
                         float neg1) {
    constexpr int NC = 32 * NG;                     // n columns: NG x 16 maps x (re, im)
    // staged B operands: [hi/lo][parity][n][same-parity l of the chunk]
    __shared__ __align__(16) _Float16 bst[2][2][NC][SH_BCOL];
    __shared__ double alds[SH_LCHUNK];
    __shared__ int wg_start;
    int mb, item;
    xcd_map(blockIdx.x, MB * nsb, &mb, &item);
    // the ring-pair blocks of one m run next to each other on one XCD, so the
    // prepared alm chunks and recursion coefficients of that m are read into
    // its L2 once (each 64-byte line of the synthesis layout belongs to one m)
    const int m = mb * MB + item / nsb;
    const int sb = item % nsb;
    if (m > mmax) return;
    const int g0 = blockIdx.y * 16 * NG;            // first map of the workgroup
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int r = lane & 31, h = lane >> 5;

    // this lane's ring pair
    const int pbase = (sb * SH_WAVES + w) * 64;
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
    // staging: the prepared chunk of each map group, 16-byte pieces (tid + 256 i)
    constexpr int PC = 2 * NG;                      // pieces per thread per chunk
    const int ngr = (nmaps + 15) / 16;
    uint4 pf[PC];
    auto fetch = [&](int c0) {
        const int64_t j = cofs[m] + (c0 - m) / SH_LCHUNK;
#pragma unroll
        for (int i = 0; i < PC; ++i) {
            const int q = tid + 256 * i, u = q >> 9, qq = q & 511;
            const int gi = blockIdx.y * NG + u;
            pf[i] = gi < ngr ? *reinterpret_cast<const uint4 *>(prep + ((int64_t)gi * nchunks + j) * SH_CHUNK_HALVES +
                                                                 8 * qq)
                             : make_uint4(0u, 0u, 0u, 0u);
        }
    };
    auto put = [&]() {
#pragma unroll
        for (int i = 0; i < PC; ++i) {
            const int q = tid + 256 * i, u = q >> 9, qq = q & 511;
            const int kb = qq & 3, n = (qq >> 2) & 31, par = (qq >> 7) & 1, hl = qq >> 8;
            *reinterpret_cast<uint4 *>(&bst[hl][par][32 * u + n][8 * kb]) = pf[i];
        }
    };
    // 2^-k of this lane's output column as two FP32 factors, exact while
    // |k| <= 250 (k >= -114 for finite rows; outputs of rows with k > 250
    // lie below the FP32 range whatever the scaling)
    float usc1[NG], usc2[NG];
#pragma unroll
    for (int u = 0; u < NG; ++u) {
        const int bo = g0 + 16 * u + (r >> 1);
        const int k = bo < nmaps ? kexp[(int64_t)bo * (mmax + 1) + m] : 0;
        usc1[u] = ldexpf(1.0f, -(k >> 1));
        usc2[u] = ldexpf(1.0f, (k >> 1) - k);
    }

    floatx16 De[2][NG], Do[2][NG];                  // [ring tile][n tile]
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int u = 0; u < NG; ++u) {
            De[j][u] = floatx16{0};
            Do[j][u] = floatx16{0};
        }
    // chunks start at l = m + 64 j (the prepared layout); the first one holds
    // every pair's switch-on
    const int lst = lstart <= lmax ? m + (lstart - m) / SH_LCHUNK * SH_LCHUNK : lmax + 1;
    if (lst <= lmax) fetch(lst);
    for (int c0 = lst; c0 <= lmax; c0 += SH_LCHUNK) {
        put();
        if (tid < SH_LCHUNK) alds[tid] = (c0 + tid <= lmax) ? al[c0 + tid] : 0.0;
        if (c0 + SH_LCHUNK <= lmax) fetch(c0 + SH_LCHUNK);     // in flight during this chunk
        __syncthreads();
        if (c0 + SH_LCHUNK > wstart) {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int l0 = c0 + 32 * s;
                if (l0 > lmax) break;
                // FP64 recursion of this lane's pair over l0 .. l0+31 (past lmax
                // alpha is staged as 0, so the values stay bounded: u_l = -u_{l-2})
                float v[32];
#pragma unroll
                for (int j = 0; j < 32; ++j) {
                    if ((j & 7) == 0 && l0 + j <= winj && ls == l0 + j) {
                        u2 = sd.x;
                        u1 = sd.y;
                    }
                    const double u = fma(alds[32 * s + j] * x, u1, -u2);
                    u2 = u1;
                    u1 = u;
                    v[j] = (float)u;
                }
                // split: block b, parity par holds l0 + 16 b + 2 i + par, i = 0..7
                uint32_t ah[2][2][4], alo[2][2][4];     // [par][block][dword]
#pragma unroll
                for (int par = 0; par < 2; ++par)
#pragma unroll
                    for (int bl = 0; bl < 2; ++bl)
#pragma unroll
                        for (int d = 0; d < 4; ++d) {
                            const float2v f = {v[16 * bl + 4 * d + par], v[16 * bl + 4 * d + 2 + par]};
                            const half2v hv = __builtin_convertvector(f, half2v);
                            const half2v lv = split_lo(f, hv, neg1);
                            ah[par][bl][d] = h2bits(hv);
                            alo[par][bl][d] = h2bits(lv);
                        }
                // tile operands: swap(block 0, block 1) -> (tile 0, tile 1)
                half8s Ah[2][2], Al[2][2];             // [par][tile]
#pragma unroll
                for (int par = 0; par < 2; ++par) {
                    uint32_t t0h[4], t1h[4], t0l[4], t1l[4];
#pragma unroll
                    for (int d = 0; d < 4; ++d) {
                        auto rh = __builtin_amdgcn_permlane32_swap(ah[par][0][d], ah[par][1][d], false, false);
                        auto rl = __builtin_amdgcn_permlane32_swap(alo[par][0][d], alo[par][1][d], false, false);
                        t0h[d] = rh[0];
                        t1h[d] = rh[1];
                        t0l[d] = rl[0];
                        t1l[d] = rl[1];
                    }
                    Ah[par][0] = __builtin_bit_cast(half8s, t0h);
                    Ah[par][1] = __builtin_bit_cast(half8s, t1h);
                    Al[par][0] = __builtin_bit_cast(half8s, t0l);
                    Al[par][1] = __builtin_bit_cast(half8s, t1l);
                }
#pragma unroll
                for (int u = 0; u < NG; ++u) {
                    // B: column 32 u + r, same-parity l 16 s + 8 h .. +7 of the chunk
                    const int cn = 32 * u + r;
                    const half8s be_h = *reinterpret_cast<const half8s *>(&bst[0][0][cn][16 * s + 8 * h]);
                    const half8s be_l = *reinterpret_cast<const half8s *>(&bst[1][0][cn][16 * s + 8 * h]);
                    const half8s bo_h = *reinterpret_cast<const half8s *>(&bst[0][1][cn][16 * s + 8 * h]);
                    const half8s bo_l = *reinterpret_cast<const half8s *>(&bst[1][1][cn][16 * s + 8 * h]);
#pragma unroll
                    for (int T = 0; T < 2; ++T) {
                        De[T][u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ah[0][T], be_h, De[T][u], 0, 0, 0);
                        Do[T][u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ah[1][T], bo_h, Do[T][u], 0, 0, 0);
                        De[T][u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ah[0][T], be_l, De[T][u], 0, 0, 0);
                        Do[T][u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ah[1][T], bo_l, Do[T][u], 0, 0, 0);
                        De[T][u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(Al[0][T], be_h, De[T][u], 0, 0, 0);
                        Do[T][u] = __builtin_amdgcn_mfma_f32_32x32x16_f16(Al[1][T], bo_h, Do[T][u], 0, 0, 0);
                    }
                }
            }
        }
        __syncthreads();
    }

    // D[row = ring][col = n]: this lane holds col = r (map r/2, re or im as
    // r is even or odd), rows (q%4) + 8 (q/4) + 4 h.  Ring pair p couples
    // ring p (north) and ring nr-1-p (south).  Adjacent lanes (re, im of one
    // map) trade one value per two rows, so that the even lane stores the
    // complex value of row q and the odd lane that of row q+1, 8 bytes each.
    // Rows 8 k .. 8 k + 7 (q < 4 over both h) are 8 ring pairs of this m:
    // one 64-byte line of the synthesis layout per map and hemisphere,
    // written whole by this wave.
    const bool odd = (r & 1) != 0;
#pragma unroll
    for (int u = 0; u < NG; ++u) {
        const int b = g0 + 16 * u + (r >> 1);
        const bool live = b < nmaps;
        const int64_t sbe = syn_block_elems(npairs), south = (int64_t)pairs8(npairs) * MB;
        float2 *out = reinterpret_cast<float2 *>(phase) + ((int64_t)min(b, nmaps - 1) * nmb + mb) * sbe + (m & 7) * 8;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int base = pbase + 32 * j;
#pragma unroll
            for (int q = 0; q < 16; q += 2) {
                // north (ge + go) and south (ge - go) values of rows q, q+1
                float vn[2], vs[2];
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const float ge = De[j][u][q + i], go = Do[j][u][q + i];
                    vn[i] = ((ge + go) * usc1[u]) * usc2[u];
                    vs[i] = ((ge - go) * usc1[u]) * usc2[u];
                }
                // even lane keeps row q and sends row q+1; odd keeps q+1, sends q
                const float sn = odd ? vn[0] : vn[1], ss = odd ? vs[0] : vs[1];
                const float rn_ = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(sn), 0xB1, 0xF, 0xF, false));
                const float rs_ = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(ss), 0xB1, 0xF, 0xF, false));
                const float2 cn = odd ? make_float2(rn_, vn[1]) : make_float2(vn[0], rn_);
                const float2 cs = odd ? make_float2(rs_, vs[1]) : make_float2(vs[0], rs_);
                const int row = ((q + (odd ? 1 : 0)) & 3) + 8 * (q >> 2) + 4 * h;
                const int pp = base + row;
                if (!live || pp >= npairs) continue;
                const int64_t o = (pp >> 3) * (8 * MB) + (pp & 7);
                out[o] = cn;
                if (nr - 1 - pp != pp) out[south + o] = cs;
            }
        }
    }
}

int launch_leg_synthesis_mfma(const Plan *p, int nmaps, const void *alm, int64_t astride, void *phase,
                              void *ws, int adjoint, hipStream_t s) {
    const int nsb = (p->npairs + SYN_PAIRS_WG - 1) / SYN_PAIRS_WG;
    const int nmb8 = (p->nmb + 7) / 8 * 8;
    dim3 grid(nmb8 * MB * nsb, (nmaps + 15) / 16);
    timer_mark(HBX_TIMER_LEGENDRE_SYNTHESIS, true, s);
    if (mfma_kind_of(p) == HBX_MFMA_F16X2) {
        // ws: prepared operands | k_b(m) [nmaps][mmax + 1]
        _Float16 *prep = static_cast<_Float16 *>(ws);
        const uint64_t pbytes = syn_prep_bytes(p, nmaps) - (uint64_t)nmaps * (p->mmax + 1) * sizeof(int);
        int *kexp = reinterpret_cast<int *>(static_cast<char *>(ws) + pbytes);
        hipLaunchKernelGGL(syn_prep_kernel, dim3(p->mmax + 1, (nmaps + 15) / 16), dim3(256), 0, s,
                           (const float2 *)alm, astride, p->d_rofs, p->d_sl, p->d_cofs, p->syn_chunks, p->lmax,
                           p->mmax, nmaps, adjoint, -1.0f, prep, kexp);
        // two 16-map groups per workgroup share each recursion; automatic
        // (tile 0): when the batch fills whole pairs of 16-map groups, where
        // the pair's cost (measured ~1.74 x one group's) is below two groups'
        const int ng = p->tile_syn_groups > 0 ? p->tile_syn_groups : (((nmaps + 15) / 16) % 2 == 0 ? 2 : 1);
        if (nmaps > 16 && ng == 2) {
            dim3 grid2(grid.x, (nmaps + 31) / 32);
            hipLaunchKernelGGL(leg_synthesis_h16_kernel<2>, grid2, dim3(SH_WAVES * 64), 0, s, (const float2 *)alm,
                               astride, (float *)phase, p->d_px, p->d_rofs, p->d_alpha, p->d_sl, p->d_lseed,
                               p->d_seed, p->npairs, p->nrings, p->lmax, p->mmax, p->nmb, nsb, nmaps, kexp, prep,
                               p->d_cofs, p->syn_chunks, -1.0f);
        } else {
            hipLaunchKernelGGL(leg_synthesis_h16_kernel<1>, grid, dim3(SH_WAVES * 64), 0, s, (const float2 *)alm,
                               astride, (float *)phase, p->d_px, p->d_rofs, p->d_alpha, p->d_sl, p->d_lseed,
                               p->d_seed, p->npairs, p->nrings, p->lmax, p->mmax, p->nmb, nsb, nmaps, kexp, prep,
                               p->d_cofs, p->syn_chunks, -1.0f);
        }
        timer_mark(HBX_TIMER_LEGENDRE_SYNTHESIS, false, s);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_error(e, "leg_synthesis_h16_kernel launch");
        return 0;
    }
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
//   A[row = n][k]  = S_par[pair 4 s + k][n]       (registers, whole kernel)
//   B[k][col]      = lambda_{c0+2col+par}(pair)   (shared LDS tile per chunk)
// A workgroup of 4 Q waves covers P = 256 Q ring pairs of one m.  Per 32-l
// chunk every wave first runs the FP64 recursion of 64 of the P pairs into
// the shared lambda tile; after one barrier each wave contracts one output
// tile -- (n-tile t, parity par) -- over its K-part of 256 pairs (64 MFMAs),
// so no wave waits on partial sums of another except the Q = 2 pair of
// K-halves, which add through a small LDS exchange.
// Lambda tile: [par][col][k][s] with pair q = 4 s + k, rows padded so that a
// recursion step's 64 lane writes and a quarter wave's 16-byte B reads hit
// disjoint banks.
// Lane layout of the 16x16x4 f32 instruction (checked on gfx950):
//   A: lane = row + 16 k   B: lane = col + 16 k   D: lane = col + 16 (row/4), reg = row%4
// Reference counterpart: the per-(m, l) ring sum of sht_raw.py:53 /
// sht_fallback.py:63.
// ---------------------------------------------------------------------------
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int ANA_L = 32;                 // l per chunk: 16 even + 16 odd columns
template <int Q> struct AnaGeom {
    static constexpr int W = 4 * Q;                    // waves per workgroup
    static constexpr int P = 256 * Q;                  // ring pairs per workgroup
    static constexpr int RUN = P / 4 + 16;             // stride between the k runs of a column
    static constexpr int COL = P + 68;                 // column stride (= 4 mod 64 banks)
    static constexpr int PAR = 16 * COL;
    static constexpr int TILE = 2 * PAR;               // floats
    static constexpr int XCH = 4 * 256;                // Q = 2 exchange: 4 tiles x 64 lanes x 4
    static constexpr size_t LDS = (size_t)(TILE + XCH) * sizeof(float) + 2 * ANA_L * sizeof(double) + 16;
};

template <int Q>
__global__ void __launch_bounds__(AnaGeom<Q>::W * 64)
leg_analysis_mfma_kernel(const float2 *__restrict__ phase, int64_t pstride, float2 *__restrict__ alm,
                         int64_t astride, float2 *__restrict__ partial, const double *__restrict__ px,
                         const int64_t *__restrict__ rofs, const double *__restrict__ alpha,
                         const double *__restrict__ sl, const int32_t *__restrict__ lseed,
                         const double2 *__restrict__ seed, int npairs, int nr, int lmax, int mmax, int ngroups,
                         int64_t nlm, int nmaps, int adjoint, int mbeg) {
    using G = AnaGeom<Q>;
    extern __shared__ __align__(16) float smem[];
    float *lam = smem;                                               // [TILE]
    float *xch = smem + G::TILE;                                     // [XCH]
    double *alds = reinterpret_cast<double *>(smem + G::TILE + G::XCH);   // [2][ANA_L]
    int *starts = reinterpret_cast<int *>(alds + 2 * ANA_L);         // [0]: workgroup, [1 + kh]: K-part

    int mb, item;
    xcd_map(blockIdx.x, MB * ngroups, &mb, &item);
    const int m = mb * MB + item / ngroups;
    const int grp = item % ngroups;
    if (m > mmax || m < mbeg) return;
    const int g = blockIdx.y;                         // group of 16 maps
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int t = w & 1, par = (w >> 1) & 1, kh = w >> 2;    // output tile, parity, K-part
    const int P0 = grp * G::P;                        // first ring pair of the workgroup

    // recursion: pair P0 + 64 w + lane
    const int pr = P0 + 64 * w + lane;
    double x = 0.0, u1 = 0.0, u2 = 0.0;
    double2 sd = make_double2(0.0, 0.0);
    int ls = NEVER_L;
    if (pr < npairs) {
        x = px[pr];
        ls = lseed[(int64_t)m * npairs + pr];
        if (ls <= lmax) sd = seed[(int64_t)m * npairs + pr];
    }
    const int wstart = wmin64(ls);
    const int winj = wmax64(ls <= lmax ? ls : -1);

    // A operands: S_par[pair P0 + 256 kh + 4 s + k][n = 16 t + row], k = lane/16, row = lane%16
    float Sv[64];
    {
        const int row = lane & 15, k = lane >> 4;
        const int n = 16 * t + row;
        const int b = g * 16 + (n >> 1), c = n & 1;
        const float *ph = reinterpret_cast<const float *>(phase + (int64_t)min(b, nmaps - 1) * pstride +
                                                          (int64_t)mb * nr * MB + (m & 7)) + c;
#pragma unroll
        for (int s = 0; s < 64; ++s) {
            const int q = P0 + 256 * kh + 4 * s + k;
            const int qq = min(q, npairs - 1);
            const float a = ph[(int64_t)qq * (2 * MB)];
            const float z = (nr - 1 - qq != qq) ? ph[(int64_t)(nr - 1 - qq) * (2 * MB)] : 0.0f;
            const bool ok = q < npairs && b < nmaps;
            Sv[s] = ok ? (par ? a - z : a + z) : 0.0f;
        }
    }
    if (tid < 1 + Q) starts[tid] = NEVER_L;
    __syncthreads();
    if (lane == 0) {
        atomicMin(&starts[0], wstart);
        atomicMin(&starts[1 + (w >> 2)], wstart);     // waves 4j..4j+3 compute K-part j's pairs
    }
    __syncthreads();
    const int lstart = starts[0];
    const int kstart = starts[1 + kh];

    const bool direct = (ngroups == 1);
    const int64_t i0 = alm_index(m, m, lmax) - m;       // + l
    const int64_t ro = rofs[m];
    const double *al = alpha + ro - m;                  // zero-padded past lmax
    const double *sv = sl + ro - m;
    // out(l, b): 2 maps x (re, im) per lane; b' = map, v = value pair
    auto store = [&](int l, int b, float2 v) {
        if (b >= nmaps) return;
        if (direct) {
            double f = sv[l];
            double re = (double)v.x * f, im = (double)v.y * f;
            if (adjoint) {
                if (m > 0) { re *= 2.0; im *= 2.0; }
                else im = 0.0;
            }
            alm[(int64_t)b * astride + i0 + l] = make_float2((float)re, (float)im);
        } else {
            partial[((int64_t)b * ngroups + grp) * nlm + i0 + l] = v;
        }
    };
    // coefficients below every ring pair's switch-on are zero
    const int lz = lstart > lmax ? lmax + 1 : lstart;
    for (int e = tid; e < (lz - m) * 16; e += G::W * 64) store(m + (e >> 4), g * 16 + (e & 15), make_float2(0.f, 0.f));
    if (lstart > lmax) return;

    if (tid < ANA_L) alds[tid] = al[lstart + tid];
    // alpha of the chunk after next is loaded a whole chunk before it is
    // staged, so the staging never waits on memory right before a barrier
    double anext = (tid < ANA_L) ? al[lstart + ANA_L + tid] : 0.0;
    __syncthreads();
    // this lane's writes: pair q = 64 w + lane -> [.][.][k = q%4][s = q/4]
    const int qw = 64 * w + lane;
    float *lw = lam + (qw & 3) * G::RUN + (qw >> 2);
    // this lane's B reads: col = lane%16, k = lane/16, s = 64 kh ..
    const float *lr = lam + par * G::PAR + (lane & 15) * G::COL + (lane >> 4) * G::RUN + 64 * kh;
    int buf = 0;
    for (int c0 = lstart; c0 <= lmax; c0 += ANA_L, buf ^= 1) {
        const double *ab = alds + buf * ANA_L;
        const double anext2 = (tid < ANA_L) ? al[c0 + 2 * ANA_L + tid] : 0.0;   // in flight a chunk
        // 1. recursion of this lane's pair over l = c0 .. c0+31 into the tile
#pragma unroll
        for (int tt = 0; tt < ANA_L; ++tt) {
            if ((tt & 7) == 0 && c0 <= winj && ls == c0 + tt) {
                u2 = sd.x;
                u1 = sd.y;
            }
            const double u = fma(ab[tt] * x, u1, -u2);
            u2 = u1;
            u1 = u;
            lw[(tt & 1) * G::PAR + (tt >> 1) * G::COL] = (float)u;
        }
        if (tid < ANA_L) alds[(buf ^ 1) * ANA_L + tid] = anext;
        anext = anext2;
        __syncthreads();
        // 2. this wave's output tile over its 256 pairs (skipped before any
        // switches on); four chains of 64 pairs keep the FP32 sums short
        floatx4 D0 = floatx4{0.f, 0.f, 0.f, 0.f}, D1 = D0, D2 = D0, D3 = D0;
        if (c0 + ANA_L > kstart) {
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const float4 bq = *reinterpret_cast<const float4 *>(lr + 4 * j);
                floatx4 &Dj = (j < 4) ? D0 : (j < 8) ? D1 : (j < 12) ? D2 : D3;
                Dj = __builtin_amdgcn_mfma_f32_16x16x4f32(Sv[4 * j], bq.x, Dj, 0, 0, 0);
                Dj = __builtin_amdgcn_mfma_f32_16x16x4f32(Sv[4 * j + 1], bq.y, Dj, 0, 0, 0);
                Dj = __builtin_amdgcn_mfma_f32_16x16x4f32(Sv[4 * j + 2], bq.z, Dj, 0, 0, 0);
                Dj = __builtin_amdgcn_mfma_f32_16x16x4f32(Sv[4 * j + 3], bq.w, Dj, 0, 0, 0);
            }
        }
        floatx4 D = (D0 + D1) + (D2 + D3);
        // 3. K-parts: part 1 hands its tile to part 0 through LDS
        if (Q == 2 && kh == 1) {
#pragma unroll
            for (int r = 0; r < 4; ++r) xch[((t * 2 + par) * 4 + r) * 64 + lane] = D[r];
        }
        __syncthreads();
        if (kh == 0) {
            if (Q == 2) {
#pragma unroll
                for (int r = 0; r < 4; ++r) D[r] += xch[((t * 2 + par) * 4 + r) * 64 + lane];
            }
            // D[row][col]: lane = col + 16 (row/4), reg = row%4; row = n - 16 t, col -> l
            const int l = c0 + 2 * (lane & 15) + par;
            if (l <= lmax) {
                const int b0 = g * 16 + 8 * t + 2 * (lane >> 4);
                store(l, b0, make_float2(D[0], D[1]));
                store(l, b0 + 1, make_float2(D[2], D[3]));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Split-fp16 matrix-core analysis.  The same contraction as
// leg_analysis_mfma_kernel, but on v_mfma_f32_16x16x32_f16 (16x the FP32
// matrix rate) with every FP32 operand carried as a pair of fp16 values,
// v = hi + lo (hi = fp16(v), lo = fp16(v - hi): 22 significant bits), and
//   S lambda = S_hi lambda_hi + S_hi lambda_lo + S_lo lambda_hi
// (the dropped S_lo lambda_lo is ~2^-22 relative); products of fp16 values
// are exact in the FP32 accumulators.  Range: the A rows of one map (re and
// im, over the workgroup's ring pairs of one parity) are scaled by 2^k so
// their maximum sits in [2^13, 2^14); k is taken without a clamp, so every
// finite FP32 input keeps its 22-bit operands, and the outputs are unscaled
// by 2^-k in FP64 at the store (exact); the rescaled recursion values stay
// far below the fp16 maximum (|mu| < 100 up to lmax 8191).
//   A[row = n][k = pair]   = S_par(pair)[n]   (registers, hi/lo, whole kernel)
//   B[k = pair][col]       = mu_{c0+2col+par}(pair)   (LDS tile, hi/lo)
// Lane layout of the 16x16x32 f16 instruction (checked on gfx950):
//   A: lane = row + 16 (k/8), halves k%8   B: lane = col + 16 (k/8), halves k%8
//   D: lane = col + 16 (row/4), reg = row%4
// Work split: 4 Q waves = (n-tile t, parity par, K-part kh of 256 pairs); the
// first 2 Q waves also run the FP64 recursion, two adjacent ring pairs per
// lane, so the hi (and lo) halves of a lane's pairs pack into one dword LDS
// store.  K-parts 1 .. Q-1 hand their tiles to K-part 0 through LDS.
// Reference counterpart: the per-(m, l) ring sum of sht_raw.py:53 /
// sht_fallback.py:63.
// ---------------------------------------------------------------------------
typedef _Float16 half8 __attribute__((ext_vector_type(8)));

template <int Q> struct H16Geom {
    static constexpr int W = 4 * Q;               // waves per workgroup
    static constexpr int P = 256 * Q;             // ring pairs per workgroup
    static constexpr int RW = 2 * Q;              // recursion waves (128 ring pairs each)
    static constexpr int COLB = 2 * P + 16;       // bytes per tile column: P halves + a 16-byte skew
    static constexpr int PARB = 16 * COLB;
    static constexpr int TILEB = 2 * PARB;        // bytes of the hi (or lo) tile
    static constexpr int XCH = (Q > 1 ? Q - 1 : 1) * 4 * 256;   // floats: K-parts 1.. hand over to K-part 0
    static constexpr size_t LDS =
        2 * (size_t)TILEB + XCH * sizeof(float) + 2 * ANA_L * sizeof(double) + (Q + 1) * sizeof(int);
};


template <int Q>
__global__ void __launch_bounds__(H16Geom<Q>::W * 64) __attribute__((amdgpu_waves_per_eu(4)))
leg_analysis_h16_kernel(const float2 *__restrict__ phase, int64_t pstride, float2 *__restrict__ alm,
                        int64_t astride, float2 *__restrict__ partial, const double *__restrict__ px,
                        const int64_t *__restrict__ rofs, const double *__restrict__ alpha,
                        const double *__restrict__ sl, const int32_t *__restrict__ lseed,
                        const double2 *__restrict__ seed, int npairs, int nr, int lmax, int mmax, int ngroups,
                        int64_t nlm, int nmaps, int adjoint, int mbeg, float neg1, int32_t *__restrict__ lfirst) {
    using G = H16Geom<Q>;
    extern __shared__ __align__(16) unsigned char hsm[];
    unsigned char *thi = hsm, *tlo = hsm + G::TILEB;
    float *xch = reinterpret_cast<float *>(hsm + 2 * G::TILEB);
    double *alds = reinterpret_cast<double *>(xch + G::XCH);          // [2][ANA_L]
    int *starts = reinterpret_cast<int *>(alds + 2 * ANA_L);         // [0]: workgroup, [1 + kh]: K-half

    // one-dimensional grid over (m-block, ring-pair group, group of 16 maps,
    // m % 8): the 8 m of one m-block (sharing 64-byte phase lines) run next
    // to each other, and the map groups of one (m, ring-pair group) on the
    // same XCD, so its recursion rows and switch-on seeds are read into L2 once
    const int ngr = (nmaps + 15) / 16;
    int mb, item;
    xcd_map(blockIdx.x, MB * ngroups * ngr, &mb, &item);
    const int m = mb * MB + item % MB;
    const int g = (item / MB) % ngr;                  // group of 16 maps
    const int grp = item / (MB * ngr);
    if (m > mmax || m < mbeg) return;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int t = w & 1, par = (w >> 1) & 1, kh = w >> 2;    // output tile, parity, K-half
    const int P0 = grp * G::P;
    const bool rec = (w < G::RW);

    // recursion: pairs P0 + 128 w + 2 lane + h, h = 0, 1 (recursion waves only)
    double x[2] = {0.0, 0.0}, u1[2] = {0.0, 0.0}, u2[2] = {0.0, 0.0};
    double2 sd[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
    int ls[2] = {NEVER_L, NEVER_L};
    if (rec) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int pr = P0 + 128 * w + 2 * lane + h;
            if (pr < npairs) {
                x[h] = px[pr];
                ls[h] = lseed[(int64_t)m * npairs + pr];
                if (ls[h] <= lmax) sd[h] = seed[(int64_t)m * npairs + pr];
            }
        }
    }
    const int wstart = wmin64(min(ls[0], ls[1]));
    const int winj = wmax64(max(ls[0] <= lmax ? ls[0] : -1, ls[1] <= lmax ? ls[1] : -1));

    // A operands: S_par[pair P0 + 256 kh + 32 s + 8 q + i][n = 16 t + row], q = lane/16, row = lane%16
    half8 Ahi[8], Alo[8];
    int kr[4];                                         // exponent k of the output rows 4 q + j
    {
        const int row = lane & 15, q = lane >> 4;
        const int n = 16 * t + row;
        const int b = g * 16 + (n >> 1), c = n & 1;
        const float *ph = reinterpret_cast<const float *>(phase + (int64_t)min(b, nmaps - 1) * pstride +
                                                          (int64_t)mb * nr * MB + (m & 7)) + c;
        // the ring value of pair (s, i) this wave loads: the north ring (parity-0
        // waves) or the south ring (parity-1 waves); the two waves of an
        // (n-tile, K-part) swap them through the still idle lambda tiles
        auto own = [&](int s, int i) -> float {
            const int qp = P0 + 256 * kh + 32 * s + 8 * q + i;
            const int qq = min(qp, npairs - 1);
            const bool ok = qp < npairs && b < nmaps && (par == 0 || nr - 1 - qq != qq);
            return ok ? ph[(int64_t)(par == 0 ? qq : nr - 1 - qq) * (2 * MB)] : 0.0f;
        };
        float *xs = reinterpret_cast<float *>(hsm);     // [W][32][64] per round of 4 K-steps
        // every load is issued before the first exchange
        float Sf[64];
#pragma unroll
        for (int s = 0; s < 8; ++s)
#pragma unroll
            for (int i = 0; i < 8; ++i) Sf[8 * s + i] = own(s, i);
        float amax = 0.0f;
#pragma unroll
        for (int rd = 0; rd < 2; ++rd) {
#pragma unroll
            for (int j = 0; j < 32; ++j) xs[(w * 32 + j) * 64 + lane] = Sf[32 * rd + j];
            __syncthreads();
#pragma unroll
            for (int j = 0; j < 32; ++j) {
                const float mine = Sf[32 * rd + j], other = xs[((w ^ 2) * 32 + j) * 64 + lane];
                const float a = par == 0 ? mine : other, z = par == 0 ? other : mine;
                Sf[32 * rd + j] = par ? a - z : a + z;
                amax = fmaxf(amax, fabsf(Sf[32 * rd + j]));
            }
            __syncthreads();
        }
        // one exponent per map (rows 2i, 2i+1 = re, im) and (n-tile, parity):
        // the K-parts (waves w & 3 + 4 j) add their products before unscaling
        amax = fmaxf(amax, __shfl_xor(amax, 1, 64));
        amax = fmaxf(amax, __shfl_xor(amax, 16, 64));
        amax = fmaxf(amax, __shfl_xor(amax, 32, 64));
        if (Q > 1) {
            float *am = xch;                            // [W][16], idle until the chunk loop
            if (lane < 16) am[w * 16 + lane] = amax;
            __syncthreads();
#pragma unroll
            for (int k2 = 0; k2 < Q; ++k2) amax = fmaxf(amax, am[((w & 3) + 4 * k2) * 16 + row]);
        }
        // map maximum scaled to [2^13, 2^14): k covers every finite FP32 row
        // (k in [-114, 162]), applied in two exact power-of-two steps
        int e = 0;
        (void)frexpf(amax, &e);
        const int k = amax > 0.0f ? 14 - e : 0;
        const int k1 = k >> 1;
        const float sc1 = ldexpf(1.0f, k1), sc2 = ldexpf(1.0f, k - k1);
#pragma unroll
        for (int s = 0; s < 8; ++s)
#pragma unroll
            for (int i = 0; i < 8; i += 2) {
                const float2v v = {(Sf[8 * s + i] * sc1) * sc2, (Sf[8 * s + i + 1] * sc1) * sc2};
                const half2v hv = __builtin_convertvector(v, half2v);
                const half2v lv = split_lo(v, hv, neg1);
                Ahi[s][i] = hv[0];
                Ahi[s][i + 1] = hv[1];
                Alo[s][i] = lv[0];
                Alo[s][i + 1] = lv[1];
            }
#pragma unroll
        for (int j = 0; j < 4; ++j) kr[j] = __shfl(k, 4 * q + j, 64);
    }
    if (tid < 1 + Q) starts[tid] = NEVER_L;
    __syncthreads();
    if (rec && lane == 0) {
        atomicMin(&starts[0], wstart);
        atomicMin(&starts[1 + (w >> 1)], wstart);     // recursion waves 2j, 2j+1 cover K-half j
    }
    __syncthreads();
    const int lstart = starts[0];
    const int kstart = starts[1 + kh];

    const bool direct = (ngroups == 1);
    const int64_t i0 = alm_index(m, m, lmax) - m;       // + l
    const int64_t ro = rofs[m];
    const double *al = alpha + ro - m;                  // zero-padded past lmax
    const double *sv = sl + ro - m;
    // v: (re, im) of one map scaled by 2^k (the map's k); unscaled in FP64
    auto store = [&](int l, int b, float2 v, int k) {
        if (b >= nmaps) return;
        if (direct) {
            double f = ldexp(sv[l], -k);
            double re = (double)v.x * f, im = (double)v.y * f;
            if (adjoint) {
                if (m > 0) { re *= 2.0; im *= 2.0; }
                else im = 0.0;
            }
            alm[(int64_t)b * astride + i0 + l] = make_float2((float)re, (float)im);
        } else {
            partial[((int64_t)b * ngroups + grp) * nlm + i0 + l] =
                make_float2((float)ldexp((double)v.x, -k), (float)ldexp((double)v.y, -k));
        }
    };
    // coefficients below every ring pair's switch-on are zero: written to alm
    // by a single group; partial sums record where theirs start instead
    const int lz = lstart > lmax ? lmax + 1 : lstart;
    if (direct)
        for (int e = tid; e < (lz - m) * 16; e += G::W * 64)
            store(m + (e >> 4), g * 16 + (e & 15), make_float2(0.f, 0.f), 0);
    else if (g == 0 && tid == 0)
        lfirst[(int64_t)m * ngroups + grp] = lz;
    if (lstart > lmax) return;

    // this lane's recursion stores: byte 2 (128 w + 2 lane) of a column
    const int wofs = 256 * w + 4 * lane;
    // this lane's B reads: column lane%16, halves 256 kh + 32 s + 8 q
    const int rofs16 = par * G::PARB + (lane & 15) * G::COLB + 2 * (256 * kh + 8 * (lane >> 4));
    // alpha_l is wave-uniform: two banks of 16 scalar registers, each loaded
    // 16 steps before its first use (a load used at once would stall the
    // wave; broadcast LDS reads would take LDS bandwidth from the operands)
    const int lst = __builtin_amdgcn_readfirstlane(lstart);
    double aA[16], aB[16];
    auto load16 = [&](double *dst, int l0) {
#pragma unroll
        for (int j = 0; j < 16; ++j) dst[j] = al[l0 + j];
    };
    // one recursion step of this lane's two pairs at l = c0 + tt into the
    // hi/lo tiles; switch-on seeds are injected at l = m + 8 j
    auto step = [&](int c0, int tt, double a) {
        if ((tt & 7) == 0 && c0 <= winj) {
#pragma unroll
            for (int h = 0; h < 2; ++h)
                if (ls[h] == c0 + tt) {
                    u2[h] = sd[h].x;
                    u1[h] = sd[h].y;
                }
        }
        float2v f;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const double u = fma(a * x[h], u1[h], -u2[h]);
            u2[h] = u1[h];
            u1[h] = u;
            f[h] = (float)u;
        }
        // packed conversions (v_cvt_pk_f16_f32, round to nearest even)
        const half2v hh = __builtin_convertvector(f, half2v);
        const half2v hl = split_lo(f, hh, neg1);
        const int o = (tt & 1) * G::PARB + (tt >> 1) * G::COLB + wofs;
        *reinterpret_cast<uint32_t *>(thi + o) = pack_h2(hh);
        *reinterpret_cast<uint32_t *>(tlo + o) = pack_h2(hl);
    };
    // 1. recursion over l = c0 .. c0+31 (aA: steps 0..15 of this chunk on
    // entry, of the next chunk on exit)
    if (rec) load16(aA, lst);
    for (int c0 = lst; c0 <= lmax; c0 += ANA_L) {
        if (rec) {
            load16(aB, c0 + 16);
#pragma unroll
            for (int tt = 0; tt < 16; ++tt) step(c0, tt, aA[tt]);
            load16(aA, c0 + ANA_L);
#pragma unroll
            for (int tt = 16; tt < ANA_L; ++tt) step(c0, tt, aB[tt - 16]);
        }
        __syncthreads();
        // 2. this wave's output tile over its 256 pairs (skipped before any switches on)
        floatx4 D0 = floatx4{0.f, 0.f, 0.f, 0.f}, D1 = D0;
        if (c0 + ANA_L > kstart) {
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const half8 bh = *reinterpret_cast<const half8 *>(thi + rofs16 + 64 * s);
                const half8 bl = *reinterpret_cast<const half8 *>(tlo + rofs16 + 64 * s);
                D0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ahi[s], bh, D0, 0, 0, 0);
                D1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ahi[s], bl, D1, 0, 0, 0);
                D1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Alo[s], bh, D1, 0, 0, 0);
            }
        }
        floatx4 D = D0 + D1;
        // 3. K-parts 1 .. Q-1 hand their tiles to K-part 0 through LDS (same row scales)
        if (Q > 1 && kh > 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) xch[((((kh - 1) * 2 + t) * 2 + par) * 4 + r) * 64 + lane] = D[r];
        }
        __syncthreads();
        if (kh == 0) {
#pragma unroll
            for (int k2 = 1; k2 < Q; ++k2)
#pragma unroll
                for (int r = 0; r < 4; ++r) D[r] += xch[((((k2 - 1) * 2 + t) * 2 + par) * 4 + r) * 64 + lane];
            // D[row][col]: lane = col + 16 (row/4), reg = row%4; row = n - 16 t, col -> l
            // (rows 4q, 4q+1 are re, im of one map and share its scale)
            const int l = c0 + 2 * (lane & 15) + par;
            if (l <= lmax) {
                const int b0 = g * 16 + 8 * t + 2 * (lane >> 4);
                store(l, b0, make_float2(D[0], D[1]), kr[0]);
                store(l, b0 + 1, make_float2(D[2], D[3]), kr[2]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Split-fp16 analysis with two 16-map groups per workgroup (32 maps; tile
// mfma_groups = 2): the contraction of leg_analysis_h16_kernel<Q> (256 Q
// ring pairs of one m, Q K-parts of 256) for both groups, 8 Q waves = (map
// group u, n-tile t, parity par, K-part kh), sharing one lambda tile per
// 32-l chunk, so every FP64 recursion step feeds twice the matrix work.  The
// recursion runs on waves 0 .. 4Q-1, one ring pair per lane; adjacent
// lanes trade their fp16 (hi, lo) halves with one DPP move, and the even lane
// stores the two pairs' hi halves, the odd lane their lo halves, as one
// dword each (the tile layout of leg_analysis_h16_kernel<Q>).  The A
// operands of each wave are loaded and scaled as there (the north/south
// exchange in four rounds of 16 K-steps, so that it fits in the idle tiles).
// Q = 2: one 16-wave workgroup per CU; Q = 1: 8-wave workgroups over 256
// ring pairs, two per CU.
// ---------------------------------------------------------------------------
template <int Q> struct H16x2Geom {
    static constexpr int W = 8 * Q;               // waves per workgroup
    static constexpr int P = 256 * Q;             // ring pairs per workgroup
    static constexpr int RW = 4 * Q;              // recursion waves (64 ring pairs each)
    static constexpr int COLB = H16Geom<Q>::COLB, PARB = H16Geom<Q>::PARB, TILEB = H16Geom<Q>::TILEB;
    // floats: the K-part 1 tiles of the 8 (u, t, par) (Q = 2), and the
    // per-wave row maxima before the chunk loop
    static constexpr int XCH = Q > 1 ? 8 * 256 : W * 16;
    static constexpr size_t LDS = 2 * (size_t)TILEB + XCH * sizeof(float) + 4 * sizeof(int);
};
static_assert((size_t)H16x2Geom<2>::W * 16 * 64 * sizeof(float) <= 2 * (size_t)H16x2Geom<2>::TILEB &&
                  (size_t)H16x2Geom<1>::W * 16 * 64 * sizeof(float) <= 2 * (size_t)H16x2Geom<1>::TILEB,
              "the A-operand exchange fits in the lambda tiles");

template <int Q>
__global__ void __launch_bounds__(H16x2Geom<Q>::W * 64) __attribute__((amdgpu_waves_per_eu(4)))
leg_analysis_h16x2_kernel(const float2 *__restrict__ phase, int64_t pstride, float2 *__restrict__ alm,
                          int64_t astride, float2 *__restrict__ partial, const double *__restrict__ px,
                          const int64_t *__restrict__ rofs, const double *__restrict__ alpha,
                          const double *__restrict__ sl, const int32_t *__restrict__ lseed,
                          const double2 *__restrict__ seed, int npairs, int nr, int lmax, int mmax, int ngroups,
                          int64_t nlm, int nmaps, int adjoint, int mbeg, float neg1, int32_t *__restrict__ lfirst) {
    using G = H16x2Geom<Q>;
    extern __shared__ __align__(16) unsigned char hsm[];
    unsigned char *thi = hsm, *tlo = hsm + G::TILEB;
    float *xch = reinterpret_cast<float *>(hsm + 2 * G::TILEB);
    int *starts = reinterpret_cast<int *>(xch + G::XCH);          // [0]: workgroup, [1 + kh]: K-half

    int mb, item;
    xcd_map(blockIdx.x, MB * ngroups, &mb, &item);
    // the 8 m of one m-block (sharing 64-byte phase lines) run next to each other
    const int m = mb * MB + item % MB;
    const int grp = item / MB;
    if (m > mmax || m < mbeg) return;
    const int g = blockIdx.y;                         // group of 32 maps
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int t = w & 1, par = (w >> 1) & 1, kh = (w >> 2) & (Q - 1), u = w / (4 * Q);
    const int P0 = grp * G::P;
    const bool rec = (w < G::RW);

    // recursion: pair P0 + 64 w + lane (recursion waves only)
    double x = 0.0, u1 = 0.0, u2 = 0.0;
    double2 sd = make_double2(0.0, 0.0);
    int ls = NEVER_L;
    if (rec) {
        const int pr = P0 + 64 * w + lane;
        if (pr < npairs) {
            x = px[pr];
            ls = lseed[(int64_t)m * npairs + pr];
            if (ls <= lmax) sd = seed[(int64_t)m * npairs + pr];
        }
    }
    const int wstart = wmin64(ls);
    const int winj = wmax64(ls <= lmax ? ls : -1);

    // A operands: S_par[pair P0 + 256 kh + 32 s + 8 q + i][n = 16 t + row] of map group u
    half8 Ahi[8], Alo[8];
    int kr[4];                                         // exponent k of the output rows 4 q + j
    {
        const int row = lane & 15, q = lane >> 4;
        const int n = 16 * t + row;
        const int b = g * 32 + 16 * u + (n >> 1), c = n & 1;
        const float *ph = reinterpret_cast<const float *>(phase + (int64_t)min(b, nmaps - 1) * pstride +
                                                          (int64_t)mb * nr * MB + (m & 7)) + c;
        auto own = [&](int s, int i) -> float {
            const int qp = P0 + 256 * kh + 32 * s + 8 * q + i;
            const int qq = min(qp, npairs - 1);
            const bool ok = qp < npairs && b < nmaps && (par == 0 || nr - 1 - qq != qq);
            return ok ? ph[(int64_t)(par == 0 ? qq : nr - 1 - qq) * (2 * MB)] : 0.0f;
        };
        float *xs = reinterpret_cast<float *>(hsm);     // [W][16][64] per round of 2 K-steps
        float Sf[64];
#pragma unroll
        for (int s = 0; s < 8; ++s)
#pragma unroll
            for (int i = 0; i < 8; ++i) Sf[8 * s + i] = own(s, i);
        float amax = 0.0f;
#pragma unroll
        for (int rd = 0; rd < 4; ++rd) {
#pragma unroll
            for (int j = 0; j < 16; ++j) xs[(w * 16 + j) * 64 + lane] = Sf[16 * rd + j];
            __syncthreads();
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const float mine = Sf[16 * rd + j], other = xs[((w ^ 2) * 16 + j) * 64 + lane];
                const float a = par == 0 ? mine : other, z = par == 0 ? other : mine;
                Sf[16 * rd + j] = par ? a - z : a + z;
                amax = fmaxf(amax, fabsf(Sf[16 * rd + j]));
            }
            __syncthreads();
        }
        // one exponent per map and (map group, n-tile, parity): the two K-halves
        // add their products before unscaling
        amax = fmaxf(amax, __shfl_xor(amax, 1, 64));
        amax = fmaxf(amax, __shfl_xor(amax, 16, 64));
        amax = fmaxf(amax, __shfl_xor(amax, 32, 64));
        float *am = xch;                                // [W][16], idle until the chunk loop
        if (lane < 16) am[w * 16 + lane] = amax;
        __syncthreads();
        float am_all = 0.0f;
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2) am_all = fmaxf(am_all, am[((w & 3) + 4 * k2 + 4 * Q * u) * 16 + row]);
        amax = am_all;
        // map maximum scaled to [2^13, 2^14), applied in two exact power-of-two steps
        int e = 0;
        (void)frexpf(amax, &e);
        const int k = amax > 0.0f ? 14 - e : 0;
        const int k1 = k >> 1;
        const float sc1 = ldexpf(1.0f, k1), sc2 = ldexpf(1.0f, k - k1);
#pragma unroll
        for (int s = 0; s < 8; ++s)
#pragma unroll
            for (int i = 0; i < 8; i += 2) {
                const float2v v = {(Sf[8 * s + i] * sc1) * sc2, (Sf[8 * s + i + 1] * sc1) * sc2};
                const half2v hv = __builtin_convertvector(v, half2v);
                const half2v lv = split_lo(v, hv, neg1);
                Ahi[s][i] = hv[0];
                Ahi[s][i + 1] = hv[1];
                Alo[s][i] = lv[0];
                Alo[s][i + 1] = lv[1];
            }
#pragma unroll
        for (int j = 0; j < 4; ++j) kr[j] = __shfl(k, 4 * q + j, 64);
    }
    if (tid < 1 + Q) starts[tid] = NEVER_L;
    __syncthreads();
    if (rec && lane == 0) {
        atomicMin(&starts[0], wstart);
        atomicMin(&starts[1 + (w >> 2)], wstart);     // recursion waves 4j .. 4j+3 cover K-part j
    }
    __syncthreads();
    const int lstart = starts[0];
    const int kstart = starts[1 + kh];

    const bool direct = (ngroups == 1);
    const int64_t i0 = alm_index(m, m, lmax) - m;       // + l
    const int64_t ro = rofs[m];
    const double *al = alpha + ro - m;                  // zero-padded past lmax
    const double *sv = sl + ro - m;
    // v: (re, im) of one map scaled by 2^k (the map's k); unscaled in FP64
    auto store = [&](int l, int b, float2 v, int k) {
        if (b >= nmaps) return;
        if (direct) {
            double f = ldexp(sv[l], -k);
            double re = (double)v.x * f, im = (double)v.y * f;
            if (adjoint) {
                if (m > 0) { re *= 2.0; im *= 2.0; }
                else im = 0.0;
            }
            alm[(int64_t)b * astride + i0 + l] = make_float2((float)re, (float)im);
        } else {
            partial[((int64_t)b * ngroups + grp) * nlm + i0 + l] =
                make_float2((float)ldexp((double)v.x, -k), (float)ldexp((double)v.y, -k));
        }
    };
    // coefficients below every ring pair's switch-on are zero: written to alm
    // by a single group; partial sums record where theirs start instead
    const int lz = lstart > lmax ? lmax + 1 : lstart;
    if (direct)
        for (int e = tid; e < (lz - m) * 32; e += G::W * 64)
            store(m + (e >> 5), g * 32 + (e & 31), make_float2(0.f, 0.f), 0);
    else if (g == 0 && tid == 0)
        lfirst[(int64_t)m * ngroups + grp] = lz;
    if (lstart > lmax) return;

    // this lane's recursion store: the dword of its lane pair in a column, in
    // the hi tile (even lanes) or the lo tile (odd lanes)
    unsigned char *wbase = ((lane & 1) ? tlo : thi) + 128 * w + 4 * (lane >> 1);
    // v_perm_b32 selectors, sources (other, mine): even lanes take the two hi
    // halves (mine, other), odd lanes the two lo halves (other, mine)
    const uint32_t psel = (lane & 1) ? 0x03020706u : 0x05040100u;
    // this lane's B reads: column lane%16, halves 256 kh + 32 s + 8 q
    const int rofs16 = par * G::PARB + (lane & 15) * G::COLB + 2 * (256 * kh + 8 * (lane >> 4));
    // alpha_l is wave-uniform: two banks of 16 scalar registers, each loaded
    // 16 steps before its first use
    const int lst = __builtin_amdgcn_readfirstlane(lstart);
    double aA[16], aB[16];
    auto load16 = [&](double *dst, int l0) {
#pragma unroll
        for (int j = 0; j < 16; ++j) dst[j] = al[l0 + j];
    };
    // two recursion steps of this lane's pair at l = c0 + tt, c0 + tt + 1 (tt
    // even) into the hi/lo tiles; switch-on seeds are injected at l = m + 8 j
    auto step2 = [&](int c0, int tt, double a0, double a1) {
        if ((tt & 7) == 0 && c0 <= winj && ls == c0 + tt) {
            u2 = sd.x;
            u1 = sd.y;
        }
        float2v f;
        {
            const double v = fma(a0 * x, u1, -u2);
            u2 = u1;
            u1 = v;
            f[0] = (float)v;
        }
        {
            const double v = fma(a1 * x, u1, -u2);
            u2 = u1;
            u1 = v;
            f[1] = (float)v;
        }
        const half2v hh = __builtin_convertvector(f, half2v);
        const half2v hl = split_lo(f, hh, neg1);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const uint32_t mine = pack_h2((half2v){hh[j], hl[j]});
            const uint32_t other = (uint32_t)__builtin_amdgcn_mov_dpp((int)mine, 0xB1, 0xF, 0xF, false);
            const uint32_t v = __builtin_amdgcn_perm(other, mine, psel);
            const int o = j * G::PARB + (tt >> 1) * G::COLB;      // l = c0 + tt + j: parity j
            *reinterpret_cast<uint32_t *>(wbase + o) = v;
        }
    };
    // 1. recursion over l = c0 .. c0+31 (aA: steps 0..15 of this chunk on
    // entry, of the next chunk on exit)
    if (rec) load16(aA, lst);
    for (int c0 = lst; c0 <= lmax; c0 += ANA_L) {
        if (rec) {
            load16(aB, c0 + 16);
#pragma unroll
            for (int tt = 0; tt < 16; tt += 2) step2(c0, tt, aA[tt], aA[tt + 1]);
            load16(aA, c0 + ANA_L);
#pragma unroll
            for (int tt = 16; tt < ANA_L; tt += 2) step2(c0, tt, aB[tt - 16], aB[tt - 15]);
        }
        __syncthreads();
        // 2. this wave's output tile over its 256 pairs (skipped before any switches on)
        floatx4 D0 = floatx4{0.f, 0.f, 0.f, 0.f}, D1 = D0;
        if (c0 + ANA_L > kstart) {
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const half8 bh = *reinterpret_cast<const half8 *>(thi + rofs16 + 64 * s);
                const half8 bl = *reinterpret_cast<const half8 *>(tlo + rofs16 + 64 * s);
                D0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ahi[s], bh, D0, 0, 0, 0);
                D1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ahi[s], bl, D1, 0, 0, 0);
                D1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(Alo[s], bh, D1, 0, 0, 0);
            }
        }
        floatx4 D = D0 + D1;
        // 3. K-half 1 hands its tile to K-half 0 through LDS (same row scales)
        const int xo = ((u * 2 + t) * 2 + par) * 256;
        if (Q > 1 && kh == 1) {
#pragma unroll
            for (int r = 0; r < 4; ++r) xch[xo + r * 64 + lane] = D[r];
        }
        __syncthreads();
        if (kh == 0) {
            if (Q > 1) {
#pragma unroll
                for (int r = 0; r < 4; ++r) D[r] += xch[xo + r * 64 + lane];
            }
            // D[row][col]: lane = col + 16 (row/4), reg = row%4; row = n - 16 t, col -> l
            // (rows 4q, 4q+1 are re, im of one map and share its scale)
            const int l = c0 + 2 * (lane & 15) + par;
            if (l <= lmax) {
                const int b0 = g * 32 + 16 * u + 8 * t + 2 * (lane >> 4);
                store(l, b0, make_float2(D[0], D[1]), kr[0]);
                store(l, b0 + 1, make_float2(D[2], D[3]), kr[2]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Pipelined matrix-core analysis (12-wave workgroups): the contraction of
// leg_analysis_mfma_kernel<2> -- 512 ring pairs of one m per workgroup, eight
// MFMA waves = 2 n-tiles x 2 parities x 2 K-halves of 256 pairs -- plus four
// recursion waves that run the FP64 lambda recursion of all 512 pairs (two
// per lane) one chunk ahead, into the other half of a double-buffered LDS
// tile.  One barrier per 32-l chunk: in step c the recursion waves fill chunk
// c+1 while the MFMA waves contract chunk c, so the FP64 VALU work overlaps
// the matrix cores instead of alternating with them.  The K-halves add
// through a double-buffered exchange that the first half reads after the
// barrier.  LDS: 2 lambda tiles (145 KB) + 2 exchange buffers + alpha.
// ---------------------------------------------------------------------------
struct PipeGeom {
    static constexpr int CW = 8, PW = 4, W = CW + PW;   // MFMA waves, recursion waves
    static constexpr int P = 512;                        // ring pairs per workgroup
    static constexpr int RUN = AnaGeom<2>::RUN, COL = AnaGeom<2>::COL, PAR = AnaGeom<2>::PAR;
    static constexpr int TILE = AnaGeom<2>::TILE;        // floats
    static constexpr int XCH = 4 * 256;                  // floats per exchange buffer
    static constexpr size_t LDS = (size_t)(2 * TILE + 2 * XCH) * sizeof(float) + 2 * ANA_L * sizeof(double) + 16;
};
static_assert(PipeGeom::P == AnaGeom<2>::P, "pipelined tile geometry follows the 8-wave kernel");

__global__ void __launch_bounds__(PipeGeom::W * 64)
leg_analysis_mfma_pipe_kernel(const float2 *__restrict__ phase, int64_t pstride, float2 *__restrict__ alm,
                              int64_t astride, float2 *__restrict__ partial, const double *__restrict__ px,
                              const int64_t *__restrict__ rofs, const double *__restrict__ alpha,
                              const double *__restrict__ sl, const int32_t *__restrict__ lseed,
                              const double2 *__restrict__ seed, int npairs, int nr, int lmax, int mmax, int ngroups,
                              int64_t nlm, int nmaps, int adjoint, int mbeg) {
    using G = PipeGeom;
    extern __shared__ __align__(16) float smem[];
    float *lam0 = smem;                                              // [2][TILE]
    float *xch0 = smem + 2 * G::TILE;                                // [2][XCH]
    double *alds = reinterpret_cast<double *>(smem + 2 * G::TILE + 2 * G::XCH);   // [2][ANA_L]
    int *starts = reinterpret_cast<int *>(alds + 2 * ANA_L);         // [0]: workgroup, [1 + kh]: K-half

    int mb, item;
    xcd_map(blockIdx.x, MB * ngroups, &mb, &item);
    const int m = mb * MB + item / ngroups;
    const int grp = item % ngroups;
    if (m > mmax || m < mbeg) return;
    const int g = blockIdx.y;                         // group of 16 maps
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool prod = (w >= G::CW);                   // recursion wave
    const int t = w & 1, par = (w >> 1) & 1, kh = (w >> 2) & 1;   // MFMA waves: tile, parity, K-half
    const int P0 = grp * G::P;
    const int64_t ro = rofs[m];
    const double *al = alpha + ro - m;                  // zero-padded past lmax
    const double *sv = sl + ro - m;
    const int64_t i0 = alm_index(m, m, lmax) - m;       // + l
    const bool direct = (ngroups == 1);

    auto store = [&](int l, int b, float2 v) {
        if (b >= nmaps) return;
        if (direct) {
            double f = sv[l];
            double re = (double)v.x * f, im = (double)v.y * f;
            if (adjoint) {
                if (m > 0) { re *= 2.0; im *= 2.0; }
                else im = 0.0;
            }
            alm[(int64_t)b * astride + i0 + l] = make_float2((float)re, (float)im);
        } else {
            partial[((int64_t)b * ngroups + grp) * nlm + i0 + l] = v;
        }
    };

    if (tid < 3) starts[tid] = NEVER_L;
    __syncthreads();

    if (prod) {
        // ----- recursion waves: pairs P0 + q0 (K-half 0) and P0 + q0 + 256 (K-half 1)
        const int q0 = 64 * (w - G::CW) + lane, ptid = tid - G::CW * 64;
        double x[2], u1[2] = {0.0, 0.0}, u2[2] = {0.0, 0.0};
        double2 sd[2];
        int ls[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int pr = P0 + q0 + 256 * h;
            x[h] = 0.0;
            sd[h] = make_double2(0.0, 0.0);
            ls[h] = NEVER_L;
            if (pr < npairs) {
                x[h] = px[pr];
                ls[h] = lseed[(int64_t)m * npairs + pr];
                if (ls[h] <= lmax) sd[h] = seed[(int64_t)m * npairs + pr];
            }
        }
        const int ws0 = wmin64(ls[0]), ws1 = wmin64(ls[1]);
        const int winj = wmax64(max(ls[0] <= lmax ? ls[0] : -1, ls[1] <= lmax ? ls[1] : -1));
        if (lane == 0) {
            atomicMin(&starts[0], min(ws0, ws1));
            atomicMin(&starts[1], ws0);
            atomicMin(&starts[2], ws1);
        }
        __syncthreads();
        const int lstart = starts[0];
        const int lz = lstart > lmax ? lmax + 1 : lstart;
        for (int e = tid; e < (lz - m) * 16; e += G::W * 64)
            store(m + (e >> 4), g * 16 + (e & 15), make_float2(0.f, 0.f));
        if (lstart > lmax) return;
        const int nch = (lmax - lstart) / ANA_L + 1;
        // alpha: chunks 0 and 1 staged now, chunk c + 2 staged in step c
        double anext = 0.0;
        if (ptid < ANA_L) {
            alds[ptid] = al[lstart + ptid];
            alds[ANA_L + ptid] = (nch > 1) ? al[lstart + ANA_L + ptid] : 0.0;
            anext = (nch > 2) ? al[lstart + 2 * ANA_L + ptid] : 0.0;
        }
        __syncthreads();
        float *lwb = lam0 + (q0 & 3) * G::RUN + (q0 >> 2);   // pair q0 + 256 sits 64 floats later
        auto produce = [&](int cc) {
            const int c0 = lstart + ANA_L * cc;
            const double *ab = alds + (cc & 1) * ANA_L;
            float *lw = lwb + (cc & 1) * G::TILE;
#pragma unroll
            for (int tt = 0; tt < ANA_L; ++tt) {
                if ((tt & 7) == 0 && c0 <= winj) {
#pragma unroll
                    for (int h = 0; h < 2; ++h)
                        if (ls[h] == c0 + tt) {
                            u2[h] = sd[h].x;
                            u1[h] = sd[h].y;
                        }
                }
                const double a = ab[tt];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const double u = fma(a * x[h], u1[h], -u2[h]);
                    u2[h] = u1[h];
                    u1[h] = u;
                    lw[(tt & 1) * G::PAR + (tt >> 1) * G::COL + 64 * h] = (float)u;
                }
            }
        };
        produce(0);
        __syncthreads();
        for (int c = 0; c < nch; ++c) {
            if (c + 1 < nch) produce(c + 1);
            if (ptid < ANA_L) {
                alds[(c & 1) * ANA_L + ptid] = anext;
                anext = (c + 3 < nch) ? al[lstart + (c + 3) * ANA_L + ptid] : 0.0;
            }
            __syncthreads();
        }
        return;
    }

    // ----- MFMA waves.  A operands: S_par[pair P0 + 256 kh + 4 s + k][n = 16 t + row]
    float Sv[64];
    {
        const int row = lane & 15, k = lane >> 4;
        const int n = 16 * t + row;
        const int b = g * 16 + (n >> 1), c = n & 1;
        const float *ph = reinterpret_cast<const float *>(phase + (int64_t)min(b, nmaps - 1) * pstride +
                                                          (int64_t)mb * nr * MB + (m & 7)) + c;
#pragma unroll
        for (int s = 0; s < 64; ++s) {
            const int q = P0 + 256 * kh + 4 * s + k;
            const int qq = min(q, npairs - 1);
            const float a = ph[(int64_t)qq * (2 * MB)];
            const float z = (nr - 1 - qq != qq) ? ph[(int64_t)(nr - 1 - qq) * (2 * MB)] : 0.0f;
            const bool ok = q < npairs && b < nmaps;
            Sv[s] = ok ? (par ? a - z : a + z) : 0.0f;
        }
    }
    __syncthreads();
    const int lstart = starts[0];
    const int kstart = starts[1 + kh];
    const int lz = lstart > lmax ? lmax + 1 : lstart;
    for (int e = tid; e < (lz - m) * 16; e += G::W * 64) store(m + (e >> 4), g * 16 + (e & 15), make_float2(0.f, 0.f));
    if (lstart > lmax) return;
    const int nch = (lmax - lstart) / ANA_L + 1;
    __syncthreads();                                    // alpha staged
    __syncthreads();                                    // chunk 0 in tile 0
    const int lro = par * G::PAR + (lane & 15) * G::COL + (lane >> 4) * G::RUN + 64 * kh;
    float *xw = xch0 + ((t * 2 + par) * 4) * 64 + lane;
    for (int c = 0; c < nch; ++c) {
        const int c0 = lstart + ANA_L * c;
        const float *lr = lam0 + (c & 1) * G::TILE + lro;
        floatx4 D0 = floatx4{0.f, 0.f, 0.f, 0.f}, D1 = D0, D2 = D0, D3 = D0;
        if (c0 + ANA_L > kstart) {
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const float4 bq = *reinterpret_cast<const float4 *>(lr + 4 * j);
                floatx4 &Dj = (j < 4) ? D0 : (j < 8) ? D1 : (j < 12) ? D2 : D3;
                Dj = __builtin_amdgcn_mfma_f32_16x16x4f32(Sv[4 * j], bq.x, Dj, 0, 0, 0);
                Dj = __builtin_amdgcn_mfma_f32_16x16x4f32(Sv[4 * j + 1], bq.y, Dj, 0, 0, 0);
                Dj = __builtin_amdgcn_mfma_f32_16x16x4f32(Sv[4 * j + 2], bq.z, Dj, 0, 0, 0);
                Dj = __builtin_amdgcn_mfma_f32_16x16x4f32(Sv[4 * j + 3], bq.w, Dj, 0, 0, 0);
            }
        }
        floatx4 D = (D0 + D1) + (D2 + D3);
        float *xb = xw + (c & 1) * G::XCH;
        if (kh == 1) {
#pragma unroll
            for (int r = 0; r < 4; ++r) xb[r * 64] = D[r];
        }
        __syncthreads();
        if (kh == 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) D[r] += xb[r * 64];
            // D[row][col]: lane = col + 16 (row/4), reg = row%4; row = n - 16 t, col -> l
            const int l = c0 + 2 * (lane & 15) + par;
            if (l <= lmax) {
                const int b0 = g * 16 + 8 * t + 2 * (lane >> 4);
                store(l, b0, make_float2(D[0], D[1]));
                store(l, b0 + 1, make_float2(D[2], D[3]));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// FP64 band of FP32 batches (m < M_LO) on the FP64 matrix cores,
// v_mfma_f64_16x16x4f64 (D[16x16] += A[16x4] B[4x16], FP64 throughout), with
// the same decomposition as the FP32 kernel: a workgroup of 8 waves covers
// 256 ring pairs of one m; each wave runs the recursion of 32 pairs into a
// shared FP64 lambda tile and contracts one output tile (n-tile, parity) over
// one K-half of 128 pairs; the K-halves add through LDS.
// Lane layout of the 16x16x4 f64 instruction (checked on gfx950):
//   A: lane = row + 16 k   B: lane = col + 16 k   D: lane = col + 16 (row%4), reg = row/4
// Output: band alm (FP64, scaled) for one group, else FP64 partial sums.
// ---------------------------------------------------------------------------
typedef double doublex4 __attribute__((ext_vector_type(4)));
struct BandGeom {
    static constexpr int W = 8, P = 256;
    static constexpr int RUN = P / 4 + 8;          // = 8 mod 32 doubles: the 4 write runs hit disjoint banks
    static constexpr int COL = 4 * RUN + 2;        // = 2 mod 32 doubles: quarter-wave 16-byte reads disjoint
    static constexpr int PAR = 16 * COL;
    static constexpr int TILE = 2 * PAR;           // doubles
    static constexpr int XCH = 4 * 256;            // doubles
    static constexpr size_t LDS = (size_t)(TILE + XCH + 2 * ANA_L) * sizeof(double) + 16;
};

__global__ void __launch_bounds__(BandGeom::W * 64)
band_analysis_mfma_kernel(const double2 *__restrict__ band, int64_t bstride, double2 *__restrict__ out,
                          double2 *__restrict__ partial, const double *__restrict__ px,
                          const int64_t *__restrict__ rofs, const double *__restrict__ alpha,
                          const double *__restrict__ sl, const int32_t *__restrict__ lseed,
                          const double2 *__restrict__ seed, int npairs, int nr, int lmax, int mtop, int ngroups,
                          int64_t nlmb, int nmaps, int adjoint) {
    using G = BandGeom;
    extern __shared__ __align__(16) double dsm[];
    double *lam = dsm;
    double *xch = dsm + G::TILE;
    double *alds = dsm + G::TILE + G::XCH;                        // [2][ANA_L]
    int *starts = reinterpret_cast<int *>(alds + 2 * ANA_L);      // [0]: workgroup, [1 + kh]

    const int m = blockIdx.x / ngroups, grp = blockIdx.x % ngroups;
    if (m > mtop) return;
    const int g = blockIdx.y;
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int t = w & 1, par = (w >> 1) & 1, kh = w >> 2;
    const int P0 = grp * G::P;

    // recursion: pair P0 + 32 w + lane % 32 (the upper half-wave repeats it)
    const int qw = 32 * w + (lane & 31);
    const int pr = P0 + qw;
    double x = 0.0, u1 = 0.0, u2 = 0.0;
    double2 sd = make_double2(0.0, 0.0);
    int ls = NEVER_L;
    if (pr < npairs) {
        x = px[pr];
        ls = lseed[(int64_t)m * npairs + pr];
        if (ls <= lmax) sd = seed[(int64_t)m * npairs + pr];
    }
    const int wstart = wmin64(ls);
    const int winj = wmax64(ls <= lmax ? ls : -1);

    // A operands: S_par[pair P0 + 128 kh + 4 s + k][n = 16 t + row]
    double Sv[32];
    {
        const int row = lane & 15, k = lane >> 4;
        const int n = 16 * t + row;
        const int b = g * 16 + (n >> 1), c = n & 1;
        const double *ph = reinterpret_cast<const double *>(band + (int64_t)min(b, nmaps - 1) * bstride + m) + c;
#pragma unroll
        for (int s = 0; s < 32; ++s) {
            const int q = P0 + 128 * kh + 4 * s + k;
            const int qq = min(q, npairs - 1);
            const double a = ph[(int64_t)qq * (2 * MB)];
            const double z = (nr - 1 - qq != qq) ? ph[(int64_t)(nr - 1 - qq) * (2 * MB)] : 0.0;
            const bool ok = q < npairs && b < nmaps;
            Sv[s] = ok ? (par ? a - z : a + z) : 0.0;
        }
    }
    if (tid < 3) starts[tid] = NEVER_L;
    __syncthreads();
    if (lane == 0) {
        atomicMin(&starts[0], wstart);
        atomicMin(&starts[1 + (w >> 2)], wstart);     // waves 4j..4j+3 run K-half j's pairs
    }
    __syncthreads();
    const int lstart = starts[0];
    const int kstart = starts[1 + kh];

    const bool direct = (ngroups == 1);
    const int64_t i0 = alm_index(m, m, lmax) - m;
    const int64_t ro = rofs[m];
    const double *al = alpha + ro - m;
    const double *sv = sl + ro - m;
    auto store = [&](int l, int b, int c, double v) {
        if (b >= nmaps) return;
        if (direct) {
            v *= sv[l];
            if (adjoint) {
                if (m > 0) v *= 2.0;
                else if (c == 1) v = 0.0;
            }
            reinterpret_cast<double *>(out + (int64_t)b * nlmb + i0 + l)[c] = v;
        } else {
            reinterpret_cast<double *>(partial + ((int64_t)b * ngroups + grp) * nlmb + i0 + l)[c] = v;
        }
    };
    const int lz = lstart > lmax ? lmax + 1 : lstart;
    for (int e = tid; e < (lz - m) * 32; e += G::W * 64) store(m + (e >> 5), g * 16 + ((e & 31) >> 1), e & 1, 0.0);
    if (lstart > lmax) return;

    if (tid < ANA_L) alds[tid] = al[lstart + tid];
    double anext = (tid < ANA_L) ? al[lstart + ANA_L + tid] : 0.0;
    __syncthreads();
    double *lw = lam + (qw & 3) * G::RUN + (qw >> 2);
    const double *lr = lam + par * G::PAR + (lane & 15) * G::COL + (lane >> 4) * G::RUN + 32 * kh;
    int buf = 0;
    for (int c0 = lstart; c0 <= lmax; c0 += ANA_L, buf ^= 1) {
        const double *ab = alds + buf * ANA_L;
        const double anext2 = (tid < ANA_L) ? al[c0 + 2 * ANA_L + tid] : 0.0;
#pragma unroll
        for (int tt = 0; tt < ANA_L; ++tt) {
            if ((tt & 7) == 0 && c0 <= winj && ls == c0 + tt) {
                u2 = sd.x;
                u1 = sd.y;
            }
            const double u = fma(ab[tt] * x, u1, -u2);
            u2 = u1;
            u1 = u;
            if (lane < 32) lw[(tt & 1) * G::PAR + (tt >> 1) * G::COL] = u;
        }
        if (tid < ANA_L) alds[(buf ^ 1) * ANA_L + tid] = anext;
        anext = anext2;
        __syncthreads();
        doublex4 D0 = {0.0, 0.0, 0.0, 0.0}, D1 = D0;
        if (c0 + ANA_L > kstart) {
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const double2 bq = *reinterpret_cast<const double2 *>(lr + 2 * j);
                D0 = __builtin_amdgcn_mfma_f64_16x16x4f64(Sv[2 * j], bq.x, D0, 0, 0, 0);
                D1 = __builtin_amdgcn_mfma_f64_16x16x4f64(Sv[2 * j + 1], bq.y, D1, 0, 0, 0);
            }
        }
        doublex4 D = D0 + D1;
        if (kh == 1) {
#pragma unroll
            for (int r = 0; r < 4; ++r) xch[((t * 2 + par) * 4 + r) * 64 + lane] = D[r];
        }
        __syncthreads();
        if (kh == 0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) D[r] += xch[((t * 2 + par) * 4 + r) * 64 + lane];
            // D[row][col]: lane = col + 16 (row%4), reg = row/4; n = 16 t + row
            const int l = c0 + 2 * (lane & 15) + par;
            if (l <= lmax) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int n = 16 * t + (lane >> 4) + 4 * r;
                    store(l, g * 16 + (n >> 1), n & 1, D[r]);
                }
            }
        }
    }
}

int band_mfma_groups(const Plan *p) { return (p->npairs + BandGeom::P - 1) / BandGeom::P; }

int launch_band_analysis_mfma(const Plan *p, int nmaps, const void *band, void *band_alm, void *band_partial,
                              int adjoint, hipStream_t s) {
    const int mtop = std::min(M_LO - 1, p->mmax);
    const int G = band_mfma_groups(p);
    const int64_t nb = band_nlm(p);
    dim3 grid((mtop + 1) * G, (nmaps + 15) / 16);
    hipLaunchKernelGGL(band_analysis_mfma_kernel, grid, dim3(BandGeom::W * 64), BandGeom::LDS, s,
                       (const double2 *)band, band_phase_elems(p), (double2 *)band_alm, (double2 *)band_partial,
                       p->d_px, p->d_rofs, p->d_alpha, p->d_sl, p->d_lseed, p->d_seed, p->npairs, p->nrings,
                       p->lmax, mtop, G, nb, nmaps, adjoint);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "band_analysis_mfma_kernel launch");
    if (G > 1) return launch_group_reduce(p, nmaps, band_partial, band_alm, nb, G, nb, 0, mtop, true, adjoint, s);
    return 0;
}

// waves per workgroup: the plan's tile (4 or 8), else 8 (4 for small
// geometries, whose ring pairs would leave a larger workgroup idle)
// operand format of the matrix-core kernels (split fp16 unless set)
int mfma_kind_of(const Plan *p) { return p->tile_mfma_kind > 0 ? p->tile_mfma_kind : HBX_MFMA_F16X2; }

// waves per matrix-core analysis workgroup: 4, 8 or (split fp16) 16, 64
// ring pairs per wave; or 12 (FP32 only: the pipelined kernel, 512 ring
// pairs, 8 MFMA + 4 recursion waves)
int mfma_analysis_waves(const Plan *p) {
    if (p->tile_mfma_waves > 0) return p->tile_mfma_waves;
    return p->npairs <= 256 ? 4 : 8;
}

int mfma_analysis_groups(const Plan *p) {
    const int W = mfma_analysis_waves(p);
    const int pairs_wg = W == 12 ? PipeGeom::P : W * 64;
    return (p->npairs + pairs_wg - 1) / pairs_wg;
}

int launch_leg_analysis_mfma(const Plan *p, int nmaps, const void *phase, void *alm, int64_t astride, void *partial,
                             int32_t *lfirst, int adjoint, hipStream_t s) {
    const int mbeg = std::min(M_LO, p->mmax + 1);    // m < M_LO run in the FP64 band
    if (mbeg > p->mmax) return 0;
    const int W = mfma_analysis_waves(p);
    const int G = mfma_analysis_groups(p);
    const int nmb8 = (p->nmb + 7) / 8 * 8;
    dim3 grid(nmb8 * MB * G, (nmaps + 15) / 16);
    const int64_t pstride = phase_elems_per_map(p);
    timer_mark(HBX_TIMER_LEGENDRE_ANALYSIS, true, s);
#define HBX_AM(QV)                                                                                              \
    hipLaunchKernelGGL(leg_analysis_mfma_kernel<QV>, grid, dim3(AnaGeom<QV>::W * 64), AnaGeom<QV>::LDS, s,     \
                       (const float2 *)phase, pstride, (float2 *)alm, astride, (float2 *)partial, p->d_px,       \
                       p->d_rofs, p->d_alpha, p->d_sl, p->d_lseed, p->d_seed, p->npairs, p->nrings, p->lmax,     \
                       p->mmax, G, p->nlm, nmaps, adjoint, mbeg)
    if (mfma_kind_of(p) == HBX_MFMA_F16X2) {
        const dim3 gridh(grid.x * grid.y);            // map groups folded into x (see the kernel)
#define HBX_AH(QV)                                                                                              \
    hipLaunchKernelGGL(leg_analysis_h16_kernel<QV>, gridh, dim3(H16Geom<QV>::W * 64), H16Geom<QV>::LDS, s,     \
                       (const float2 *)phase, pstride, (float2 *)alm, astride, (float2 *)partial, p->d_px,       \
                       p->d_rofs, p->d_alpha, p->d_sl, p->d_lseed, p->d_seed, p->npairs, p->nrings, p->lmax,     \
                       p->mmax, G, p->nlm, nmaps, adjoint, mbeg, -1.0f, lfirst)
        // two 16-map groups per workgroup (tile mfma_groups = 2; results
        // bit-identical to one group): automatic only where one 256-pair
        // workgroup covers every ring pair (nside <= 128) and the 16-map
        // groups pair up (map2alm of 32 maps at nside 128: 0.128 against
        // 0.137 ms).  Over several ring-pair groups it does not pay: the
        // 8-wave geometry's single 16-wave workgroup per CU alternates
        // recursion and matrix phases with nothing to overlap them (nside
        // 2048, 32 maps: 24.9 against 21.3 ms), and the 256-pair geometry
        // halves the recursion for a 6 % shorter kernel (21.1 against 22.4
        // ms) but doubles the partial sums (map2alm 38.1 against 35.6 ms).
#define HBX_AH2(QV)                                                                                             \
    hipLaunchKernelGGL(leg_analysis_h16x2_kernel<QV>, grid2, dim3(H16x2Geom<QV>::W * 64), H16x2Geom<QV>::LDS, s, \
                       (const float2 *)phase, pstride, (float2 *)alm, astride, (float2 *)partial, p->d_px,       \
                       p->d_rofs, p->d_alpha, p->d_sl, p->d_lseed, p->d_seed, p->npairs, p->nrings, p->lmax,     \
                       p->mmax, G, p->nlm, nmaps, adjoint, mbeg, -1.0f, lfirst)
        const bool pair_groups = ((nmaps + 15) / 16) % 2 == 0;
        const bool two = p->tile_syn_groups == 2 || (p->tile_syn_groups == 0 && W == 4 && G == 1 && pair_groups);
        if ((W == 4 || W == 8) && nmaps > 16 && two) {
            dim3 grid2(grid.x, (nmaps + 31) / 32);
            if (W == 4) HBX_AH2(1);
            else HBX_AH2(2);
        } else switch (W) {
        case 4: HBX_AH(1); break;
        case 8: HBX_AH(2); break;
        case 16: HBX_AH(4); break;
        default: return set_error(HBX_EUNSUPPORTED, "split-fp16 analysis waves must be 4, 8 or 16");
        }
#undef HBX_AH
#undef HBX_AH2
    } else switch (W) {
    case 4: HBX_AM(1); break;
    case 8: HBX_AM(2); break;
    case 12:
        hipLaunchKernelGGL(leg_analysis_mfma_pipe_kernel, grid, dim3(PipeGeom::W * 64), PipeGeom::LDS, s,
                           (const float2 *)phase, pstride, (float2 *)alm, astride, (float2 *)partial, p->d_px,
                           p->d_rofs, p->d_alpha, p->d_sl, p->d_lseed, p->d_seed, p->npairs, p->nrings, p->lmax,
                           p->mmax, G, p->nlm, nmaps, adjoint, mbeg);
        break;
    default: return set_error(HBX_EUNSUPPORTED, "matrix-core analysis waves must be 4, 8 or 12");
    }
#undef HBX_AM
    timer_mark(HBX_TIMER_LEGENDRE_ANALYSIS, false, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "leg_analysis_mfma_kernel launch");
    if (G > 1)
        return launch_group_reduce(p, nmaps, partial, alm, astride, G, p->nlm, mbeg, p->mmax, false, adjoint, s,
                                   mfma_kind_of(p) == HBX_MFMA_F16X2 ? lfirst : nullptr);
    return 0;
}

// raise the dynamic-LDS limit of the analysis kernels (called at plan creation)
int prepare_mfma_kernels() {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(leg_analysis_mfma_kernel<1>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)AnaGeom<1>::LDS);
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(leg_analysis_mfma_kernel<2>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)AnaGeom<2>::LDS);
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(leg_analysis_h16_kernel<1>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)H16Geom<1>::LDS);
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(leg_analysis_h16_kernel<2>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)H16Geom<2>::LDS);
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(leg_analysis_h16_kernel<4>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)H16Geom<4>::LDS);
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(leg_analysis_h16x2_kernel<1>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)H16x2Geom<1>::LDS);
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(leg_analysis_h16x2_kernel<2>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)H16x2Geom<2>::LDS);
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(leg_analysis_mfma_pipe_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)PipeGeom::LDS);
    if (e == hipSuccess)
        e = hipFuncSetAttribute(reinterpret_cast<const void *>(band_analysis_mfma_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)BandGeom::LDS);
    if (e != hipSuccess) return hip_error(e, "hipFuncSetAttribute(matrix-core analysis kernels)");
    return 0;
}

}  // namespace hbx
