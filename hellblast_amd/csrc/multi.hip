// Mixed-geometry batches in persistent launches (no reference counterpart:
// the reference transforms one map per call, sht_raw.py:30-57).
//
// A batch object (hbx_multi) holds up to HBX_MULTI_MAX_GROUPS plans
// ("groups": one geometry, lmax and mmax each) and a map count per group.
// One call transforms every map of every group.  The ring stage runs in
// fused launches, one per workgroup size of the FFT size classes, each over
// the rings of every group (ring_fft.hip); each Legendre direction is ONE
// persistent launch whose workgroups pull tiles -- (group, m-block or part
// of it, ring-pair block, map block) -- from work queues until every tile
// is done, so the m-blocks of one large map and the tiles of many small
// maps spread over the whole GPU in a single kernel.  A call therefore
// issues a fixed number of launches, whatever the number of groups.
//
// Queues: the tiles are split into 8 queues by m-block (all tiles of one
// m-block, whose 8 m share the 64-byte lines of the analysis phase buffer,
// sit in one queue), and each queue is ordered by estimated cost, largest
// first (the cost counts the l steps every wave runs from its first
// switched-on ring pair).  Workgroup w starts on queue w % 8 -- consecutive
// workgroups are dispatched round-robin over the 8 XCDs, so a queue's tiles
// mostly run on one XCD and share its L2 -- and pulls from the next queue in
// turn once its own is empty.  A pull is one atomicAdd on the queue's
// counter; a workgroup leaves after a failed pull on each of the 8 queues,
// so every workgroup exits once all tiles are taken.  The counters live in
// the call's workspace and are zeroed by the call's first kernel (no memset
// node), so a call can be captured in a HIP graph and replayed.
//
// Tiles run the wave bodies of the per-plan VALU kernels (legendre_valu.h):
// the analysis with R ring-pair slots per lane (64 R ring pairs of one m and
// one map per single-wave workgroup, partial sums per ring-pair block reduced
// in a fixed order), the synthesis with the 8 m of an m-block over 64 ring
// pairs per 8-wave workgroup and up to BM maps of a group per tile sharing
// each FP64 recursion step.  The results therefore equal the per-plan VALU
// path bit for bit when it uses the same slots.  FP32 groups take m < 8 in
// FP64, as in hbx_map2alm: one band-DFT launch over every group's rings
// (the per-plan VALU band DFT) and one persistent FP64 launch over band
// tiles with min(4, slots) ring-pair slots (the per-plan VALU band), which
// takes the second set of 8 queue counters.
#include <algorithm>
#include <climits>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "hbx_device.h"
#include "hbx_internal.h"
#include "legendre_valu.h"

namespace hbx {

constexpr int MULTI_MAXG = HBX_MULTI_MAX_GROUPS;
constexpr int NQ = 8;                       // work queues (one per XCD)
constexpr uint64_t HDR_CALL = 256;          // workspace: counters at 0, per-call pointers at 256
constexpr uint64_t HDR_BYTES = 1024;

// one group's tables and workspace offsets (device copy made at creation)
struct MultiGeomDev {
    const int32_t *pair_n, *pair_s, *lseed;
    const double *px, *alpha, *sl;
    const int64_t *rofs;
    const double2 *seed;
    int npairs, nr, lmax, mmax, ngr, nmaps;
    int64_t nlm, pstride, nrow;
    uint64_t phase_off, aux_off;   // bytes into the workspace
    uint64_t out_off;              // analysis result: 0 = the call's alm, else rows of nlm in the workspace
};

// analysis tile: m0 .. m0 + mn - 1 (one m-block), ring-pair block rb of 64 R
// pairs, map b0 (nb = 1).  Synthesis tile: the m-block starting at m0
// (mn = 8), pair block rb of 64 pairs, maps b0 .. b0 + nb - 1.
struct MultiTile {
    int32_t g, m0, mn, rb, b0, nb;
};

// one (group, m, map) row of the reduction / pre-scaling passes
struct MultiRow {
    int32_t g, m, b, pad;
};

// per-call pointers, copied into the workspace by the call's first kernel
struct MultiCall {
    void *alm[MULTI_MAXG];
    int64_t astride[MULTI_MAXG];
    void *maps[MULTI_MAXG];
    int64_t mstride[MULTI_MAXG];
};
static_assert(sizeof(MultiCall) <= HDR_BYTES - HDR_CALL, "per-call header");

__global__ void multi_setup_kernel(MultiCall c, char *ws) {
    if (threadIdx.x < 2 * NQ) reinterpret_cast<unsigned *>(ws)[threadIdx.x] = 0u;
    if (threadIdx.x == 0) *reinterpret_cast<MultiCall *>(ws + HDR_CALL) = c;
}

// next tile index for the calling thread: queues home, home+1, ... in turn;
// qi (the thread's current queue offset) only grows; -1 once all are empty.
// The home queue is pulled directly; another queue is first checked with a
// coherent load of its counter (which only grows), so a drained queue costs
// a load instead of an atomic on a counter every workgroup hits at the end
// of the launch.  Measured (16 FP32 maps at nside 64 in 16 groups / 16 at
// nside 512 / the 64-map nside 128..2048 batch, ms per call pair): atomics
// on every queue 0.86 / 2.88 / 13.8, a load before every pull 0.18 / 3.68 /
// 15.7, a load before steals only 0.27 / 2.56 / 13.6.
__device__ __forceinline__ int pull_tile(unsigned *queue, const int32_t *__restrict__ qrange, int home, int &qi) {
    while (qi < NQ) {
        const int q = (home + qi) & (NQ - 1);
        const unsigned n = (unsigned)(qrange[NQ + q] - qrange[q]);
        if (qi == 0 || __hip_atomic_load(queue + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < n) {
            const int t = qrange[q] + (int)atomicAdd(queue + q, 1u);
            if (t < qrange[NQ + q]) return t;
        }
        ++qi;
    }
    return -1;
}

template <typename T, int R>
__global__ void __launch_bounds__(64)
multi_analysis_kernel(const MultiGeomDev *__restrict__ geo, const MultiTile *__restrict__ tiles,
                      const int32_t *__restrict__ qrange, char *__restrict__ ws, int qoff) {
    __shared__ T red[red_elems<T>()];
    unsigned *queue = reinterpret_cast<unsigned *>(ws) + qoff;
    const MultiCall *call = reinterpret_cast<const MultiCall *>(ws + HDR_CALL);
    const int lane = threadIdx.x;
    const int home = blockIdx.x & (NQ - 1);
    int qi = 0;
    // one m per trip: the tile's m range is walked by this loop rather than
    // an inner one, so per-lane addresses are not kept live across it
    MultiTile tl{0, 0, 0, 0, 0, 0};
    int m = 0;
    for (;;) {
        if (m >= tl.m0 + tl.mn) {
            int t = 0;
            if (lane == 0) t = pull_tile(queue, qrange, home, qi);
            t = uniform(t);   // lane 0's value
            if (t < 0) break;
            tl = tiles[t];
            m = tl.m0;
        }
        const MultiGeomDev G = geo[tl.g];
        const c2<T> *ph = reinterpret_cast<const c2<T> *>(ws + G.phase_off) + (int64_t)tl.b0 * G.pstride;
        const bool direct = G.ngr == 1;
        c2<T> *dst = !direct   ? reinterpret_cast<c2<T> *>(ws + G.aux_off) + ((int64_t)tl.b0 * G.ngr + tl.rb) * G.nlm
                     : G.out_off ? reinterpret_cast<c2<T> *>(ws + G.out_off) + (int64_t)tl.b0 * G.nlm
                                 : reinterpret_cast<c2<T> *>(call->alm[tl.g]) + (int64_t)tl.b0 * call->astride[tl.g];
        ana_wave<T, R>(ph + (int64_t)(m >> 3) * G.nr * MB + (m & 7), dst, direct, G.pair_n, G.pair_s, G.px, G.rofs,
                       G.alpha, G.sl, G.lseed, G.seed, G.npairs, G.lmax, m, tl.rb, 0, red, lane);
        ++m;
    }
}

template <typename T>
__global__ void multi_reduce_kernel(const MultiGeomDev *__restrict__ geo, const MultiRow *__restrict__ rows,
                                    char *__restrict__ ws) {
    const MultiRow r = rows[blockIdx.x];
    const MultiGeomDev G = geo[r.g];
    const MultiCall *call = reinterpret_cast<const MultiCall *>(ws + HDR_CALL);
    const int64_t i0 = alm_index(r.m, r.m, G.lmax);
    const c2<T> *src = reinterpret_cast<const c2<T> *>(ws + G.aux_off) + (int64_t)r.b * G.ngr * G.nlm + i0;
    c2<T> *dst = (G.out_off ? reinterpret_cast<c2<T> *>(ws + G.out_off) + (int64_t)r.b * G.nlm
                            : reinterpret_cast<c2<T> *>(call->alm[r.g]) + (int64_t)r.b * call->astride[r.g]) +
                 i0;
    reduce_row<T>(src, G.nlm, G.ngr, dst, G.sl + G.rofs[r.m], G.lmax, r.m, 0, nullptr);
}

// FP32 batches: the FP64 band alm (rows m < M_LO, the first nlm entries of
// the m-major layout) of one (group, map) row rounded into the call's alm
__global__ void multi_band_convert_kernel(const MultiGeomDev *__restrict__ bgeo, const MultiRow *__restrict__ rows,
                                          char *__restrict__ ws) {
    const MultiRow r = rows[blockIdx.x];
    const MultiGeomDev G = bgeo[r.g];
    const MultiCall *call = reinterpret_cast<const MultiCall *>(ws + HDR_CALL);
    const double2 *src = reinterpret_cast<const double2 *>(ws + G.out_off) + (int64_t)r.b * G.nlm;
    float2 *dst = reinterpret_cast<float2 *>(call->alm[r.g]) + (int64_t)r.b * call->astride[r.g];
    for (int64_t i = threadIdx.x; i < G.nlm; i += blockDim.x) {
        const double2 v = src[i];
        dst[i] = make_float2((float)v.x, (float)v.y);
    }
}

template <typename T>
__global__ void multi_prescale_kernel(const MultiGeomDev *__restrict__ geo, const MultiRow *__restrict__ rows,
                                      char *__restrict__ ws) {
    const MultiRow r = rows[blockIdx.x];
    const MultiGeomDev G = geo[r.g];
    const MultiCall *call = reinterpret_cast<const MultiCall *>(ws + HDR_CALL);
    const int64_t o = G.rofs[r.m];
    const c2<T> *a =
        reinterpret_cast<const c2<T> *>(call->alm[r.g]) + (int64_t)r.b * call->astride[r.g] + alm_index(r.m, r.m, G.lmax);
    c2<T> *d = reinterpret_cast<c2<T> *>(ws + G.aux_off) + (int64_t)r.b * G.nrow + o;
    prescale_row<T>(a, d, G.sl + o, (int)(G.rofs[r.m + 1] - o), G.lmax, r.m, 0);
}

template <typename T, int NB>
__device__ __forceinline__ void syn_tile(const MultiGeomDev &G, const MultiTile &tl, char *ws, int m, int p) {
    const int64_t ro = ldc(G.rofs + m);
    const c2<T> *sc = reinterpret_cast<const c2<T> *>(ws + G.aux_off) + (int64_t)tl.b0 * G.nrow + ro - m;
    c2<T> *ph = reinterpret_cast<c2<T> *>(ws + G.phase_off) + (int64_t)tl.b0 * G.pstride +
                syn_ofs(m, 0, p < G.npairs ? p : 0, G.npairs);
    const c2<T> *ar[NB];
    c2<T> *out[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        ar[j] = sc + (int64_t)j * G.nrow;
        out[j] = ph + (int64_t)j * G.pstride;
    }
    syn_wave<T, NB>(ar, out, G.pair_n, G.pair_s, G.px, G.alpha + ro - m, G.lseed, G.seed, G.npairs, G.lmax, m, p);
}

template <typename T, int BM>
__global__ void __launch_bounds__(512)
multi_synthesis_kernel(const MultiGeomDev *__restrict__ geo, const MultiTile *__restrict__ tiles,
                       const int32_t *__restrict__ qrange, char *__restrict__ ws) {
    __shared__ int s_tile;
    unsigned *queue = reinterpret_cast<unsigned *>(ws);
    const int w = uniform(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int home = blockIdx.x & (NQ - 1);
    int qi = 0;
    for (;;) {
        if (threadIdx.x == 0) s_tile = pull_tile(queue, qrange, home, qi);
        __syncthreads();
        const int t = uniform(s_tile);
        __syncthreads();   // every wave has read s_tile before thread 0 pulls again
        if (t < 0) break;
        const MultiTile tl = tiles[t];
        const MultiGeomDev G = geo[tl.g];
        const int m = tl.m0 + w;
        if (m > G.mmax) continue;
        const int p = tl.rb * 64 + lane;
        switch (tl.nb) {
        case 1: syn_tile<T, 1>(G, tl, ws, m, p); break;
        case 2: if constexpr (BM >= 2) syn_tile<T, 2>(G, tl, ws, m, p); break;
        case 3: if constexpr (BM >= 3) syn_tile<T, 3>(G, tl, ws, m, p); break;
        case 4: if constexpr (BM >= 4) syn_tile<T, 4>(G, tl, ws, m, p); break;
        default: break;
        }
    }
}

static uint64_t align256(uint64_t v) { return (v + 255) & ~uint64_t(255); }

}  // namespace hbx

using namespace hbx;

struct hbx_multi {
    int device = 0, precision = 0, ngroups = 0;
    int slots = 0, syn_maps = 0, tiles_per_wave = 0;
    std::vector<Plan> plans;               // copies: share each plan's device tables
    std::vector<int> nmaps, ngr;
    std::vector<uint64_t> phase_off, aux_off;
    uint64_t ws_bytes = 0;
    MultiGeomDev *d_geo = nullptr;
    MultiTile *d_atiles = nullptr, *d_stiles = nullptr;
    int32_t *d_aq = nullptr, *d_sq = nullptr;      // [2 * NQ] queue ranges: begin, end
    MultiRow *d_rrows = nullptr, *d_prows = nullptr;
    int n_atiles = 0, n_stiles = 0, n_rrows = 0, n_prows = 0;
    int grid_ana = 0, grid_syn = 0;
    // ring stage: fused launches over every group's size classes
    std::vector<RingLaunch> rings;
    RingLaunch band_dft;
    // FP32 batches, FP64 band (m < M_LO): persistent tiles, reduction rows,
    // conversion rows (one per group and map)
    int band_slots = 0, grid_band = 0, n_btiles = 0, n_brrows = 0, n_bcrows = 0;
    MultiGeomDev *d_bgeo = nullptr;
    MultiTile *d_btiles = nullptr;
    int32_t *d_bq = nullptr;
    MultiRow *d_brrows = nullptr, *d_bcrows = nullptr;
    // side stream of the band chain (non-blocking, default priority) and the
    // events that fork it from and join it to the caller's stream, used
    // under side_mu from fork to join
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    mutable std::mutex side_mu;
    std::vector<void *> ptrs;
    ~hbx_multi() {
        int cur = 0;
        const bool have = hipGetDevice(&cur) == hipSuccess;
        if (have && cur != device) (void)hipSetDevice(device);
        for (void *p : ptrs) (void)hipFree(p);
        if (side) (void)hipStreamDestroy(side);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (have && cur != device) (void)hipSetDevice(cur);
    }
};

namespace hbx {

template <typename X> static int upload_vec(hbx_multi *mb, X **dst, const std::vector<X> &v) {
    *dst = nullptr;
    if (v.empty()) return HBX_OK;
    void *ptr = nullptr;
    hipError_t e = hipMalloc(&ptr, v.size() * sizeof(X));
    if (e != hipSuccess) return set_error(HBX_ENOMEM, "hipMalloc for the batch tables failed");
    mb->ptrs.push_back(ptr);
    HBX_HIP(hipMemcpy(ptr, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice));
    *dst = static_cast<X *>(ptr);
    return HBX_OK;
}

// tiles sorted into the 8 queues (by m-block), each by cost, largest first
static void sort_queues(std::vector<MultiTile> &tiles, const std::vector<double> &cost, std::vector<int32_t> &qrange) {
    std::vector<size_t> idx(tiles.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
    auto q_of = [&](size_t i) { return (tiles[i].m0 / MB + 3 * tiles[i].g) & (NQ - 1); };
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
        const int qa = q_of(a), qb = q_of(b);
        if (qa != qb) return qa < qb;
        return cost[a] > cost[b];
    });
    std::vector<MultiTile> out(tiles.size());
    qrange.assign(2 * NQ, 0);
    for (int q = 0; q < NQ; ++q) qrange[q] = qrange[NQ + q] = 0;
    std::vector<int> cnt(NQ, 0);
    for (size_t i = 0; i < idx.size(); ++i) {
        out[i] = tiles[idx[i]];
        ++cnt[q_of(idx[i])];
    }
    int o = 0;
    for (int q = 0; q < NQ; ++q) {
        qrange[q] = o;
        o += cnt[q];
        qrange[NQ + q] = o;
    }
    tiles.swap(out);
}

static int resident_blocks(const void *kernel, int threads, int *out) {
    int dev = 0, cus = 0, per = 0;
    HBX_HIP(hipGetDevice(&dev));
    HBX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HBX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0));
    int n = std::max(1, per) * std::max(1, cus);
    *out = std::max(NQ, n / NQ * NQ);
    return HBX_OK;
}

template <typename T> static const void *ana_kernel(int R) {
    switch (R) {
    case 1: return reinterpret_cast<const void *>(multi_analysis_kernel<T, 1>);
    case 2: return reinterpret_cast<const void *>(multi_analysis_kernel<T, 2>);
    case 4: return reinterpret_cast<const void *>(multi_analysis_kernel<T, 4>);
    case 8: return reinterpret_cast<const void *>(multi_analysis_kernel<T, 8>);
    default: return nullptr;
    }
}
template <typename T> static const void *syn_kernel(int BM) {
    switch (BM) {
    case 1: return reinterpret_cast<const void *>(multi_synthesis_kernel<T, 1>);
    case 2: return reinterpret_cast<const void *>(multi_synthesis_kernel<T, 2>);
    case 4: return reinterpret_cast<const void *>(multi_synthesis_kernel<T, 4>);
    default: return nullptr;
    }
}

// first switched-on l of each wave (64 consecutive ring pairs of one m): the
// l at which its recursion starts
static int wave_starts(const Plan *p, std::vector<int> &ws0) {
    const int np = p->npairs, nw = (np + 63) / 64;
    std::vector<int32_t> ls((size_t)(p->mmax + 1) * np);
    HBX_HIP(hipMemcpy(ls.data(), p->d_lseed, ls.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    ws0.assign((size_t)(p->mmax + 1) * nw, NEVER);
    for (int m = 0; m <= p->mmax; ++m)
        for (int q = 0; q < np; ++q) {
            int &v = ws0[(size_t)m * nw + q / 64];
            v = std::min(v, (int)ls[(size_t)m * np + q]);
        }
    return HBX_OK;
}

// Analysis cost per (m, ring-pair block of 64 R pairs): the l steps every
// wave of the block runs from its first switched-on l
static std::vector<double> block_costs(const Plan *p, const std::vector<int> &w0, int R) {
    const int nw = (p->npairs + 63) / 64, ngr = (p->npairs + 64 * R - 1) / (64 * R);
    std::vector<double> c((size_t)(p->mmax + 1) * ngr, 0.0);
    for (int m = 0; m <= p->mmax; ++m)
        for (int w = 0; w < nw; ++w) {
            const int s = std::max(w0[(size_t)m * nw + w], m);
            if (s <= p->lmax) c[(size_t)m * ngr + w / R] += 64.0 * (p->lmax + 1 - s);
        }
    return c;
}

// analysis tiles of one group for m in [mlo, mhi]: consecutive m of one
// m-block, about `target` l-steps each (at least one m)
static void analysis_tiles(int g, int nmaps, int ngr, int mlo, int mhi, const std::vector<double> &cm,
                           double target, std::vector<MultiTile> &at, std::vector<double> &acost) {
    for (int b = 0; b < nmaps; ++b)
        for (int r = 0; r < ngr; ++r)
            for (int m = mlo; m <= mhi;) {
                const int mend = std::min(mhi, (m / MB) * MB + MB - 1);
                MultiTile t{g, m, 0, r, b, 1};
                double c = 0.0;
                while (m <= mend && (t.mn == 0 || c + cm[(size_t)m * ngr + r] <= target)) {
                    c += cm[(size_t)m * ngr + r];
                    ++t.mn;
                    ++m;
                }
                at.push_back(t);
                acost.push_back(c);
            }
}

static MultiGeomDev geom_of(const Plan *p, int nmaps, int ngr) {
    MultiGeomDev d{};
    d.pair_n = p->d_pair_n;
    d.pair_s = p->d_pair_s;
    d.lseed = p->d_lseed;
    d.px = p->d_px;
    d.alpha = p->d_alpha;
    d.sl = p->d_sl;
    d.rofs = p->d_rofs;
    d.seed = p->d_seed;
    d.npairs = p->npairs;
    d.nr = p->nrings;
    d.lmax = p->lmax;
    d.mmax = p->mmax;
    d.ngr = ngr;
    d.nmaps = nmaps;
    d.nlm = p->nlm;
    d.pstride = phase_elems_per_map(p);
    d.nrow = p->nrow;
    return d;
}

static int build_multi(hbx_multi *mb) {
    const bool f64 = mb->precision == HBX_F64;
    const int R = mb->slots, BM = mb->syn_maps;
    // FP64 band of FP32 batches: at most 4 slots (the per-plan VALU band's
    // limit), so a plan tiled with the batch's slots runs the same band
    const int RB = std::min(4, R);
    mb->band_slots = f64 ? 0 : RB;
    const void *ka = f64 ? ana_kernel<double>(R) : ana_kernel<float>(R);
    const void *ks = f64 ? syn_kernel<double>(BM) : syn_kernel<float>(BM);
    int rc;
    if ((rc = resident_blocks(ka, 64, &mb->grid_ana))) return rc;
    if ((rc = resident_blocks(ks, 512, &mb->grid_syn))) return rc;
    if (!f64 && (rc = resident_blocks(ana_kernel<double>(RB), 64, &mb->grid_band))) return rc;

    const int G = mb->ngroups;
    const uint64_t cs = f64 ? 16 : 8;
    std::vector<MultiGeomDev> geo(G), bgeo;
    std::vector<MultiTile> at, st, bt;
    std::vector<double> acost, scost, bcost;
    std::vector<MultiRow> rrows, prows, brrows, bcrows;
    uint64_t off = HDR_BYTES;
    mb->ngr.assign(G, 1);
    mb->phase_off.assign(G, 0);
    mb->aux_off.assign(G, 0);
    std::vector<std::vector<double>> cm(G), cb(G);
    std::vector<std::vector<int>> wst(G);
    std::vector<int> ngb(G, 1);
    double total = 0.0, btotal = 0.0;
    for (int g = 0; g < G; ++g) {
        const Plan *p = &mb->plans[g];
        const int n = mb->nmaps[g];
        const int ngr = (p->npairs + 64 * R - 1) / (64 * R);
        mb->ngr[g] = ngr;
        if ((rc = wave_starts(p, wst[g]))) return rc;
        cm[g] = block_costs(p, wst[g], R);
        const int mlo = f64 ? 0 : std::min(M_LO, p->mmax + 1);
        for (int m = mlo; m <= p->mmax; ++m)
            for (int r = 0; r < ngr; ++r) total += cm[g][(size_t)m * ngr + r] * n;
        // workspace: phase buffer (+ FP32 band phase), then the aux region:
        // analysis partial sums | FP32 band alm | band partials, or the
        // synthesis scaled rows
        mb->phase_off[g] = off;
        off += align256(phase_total_bytes(p, n));
        mb->aux_off[g] = off;
        uint64_t ana = align256(ngr > 1 ? (uint64_t)n * ngr * p->nlm * cs : 0);
        geo[g] = geom_of(p, n, ngr);
        geo[g].phase_off = mb->phase_off[g];
        geo[g].aux_off = mb->aux_off[g];
        if (!f64) {
            ngb[g] = (p->npairs + 64 * RB - 1) / (64 * RB);
            cb[g] = block_costs(p, wst[g], RB);
            const int mt = std::min(M_LO - 1, p->mmax);
            for (int m = 0; m <= mt; ++m)
                for (int r = 0; r < ngb[g]; ++r) btotal += cb[g][(size_t)m * ngb[g] + r] * n;
            MultiGeomDev d = geom_of(p, n, ngb[g]);
            d.mmax = mt;
            d.nlm = band_nlm(p);
            d.pstride = band_phase_elems(p);
            d.phase_off = mb->phase_off[g] + band_phase_offset(p, n);
            d.out_off = mb->aux_off[g] + ana;                       // band alm (FP64)
            ana += align256((uint64_t)n * band_nlm(p) * 16);
            d.aux_off = mb->aux_off[g] + ana;                       // band partial sums
            ana += align256(ngb[g] > 1 ? (uint64_t)n * ngb[g] * band_nlm(p) * 16 : 0);
            bgeo.push_back(d);
        }
        const uint64_t syn = align256((uint64_t)n * p->nrow * cs);
        off += std::max(ana, syn);
    }
    mb->ws_bytes = off;
    const double target = total / ((double)mb->grid_ana * mb->tiles_per_wave);
    const double btarget = f64 ? 0.0 : btotal / ((double)mb->grid_band * mb->tiles_per_wave);
    for (int g = 0; g < G; ++g) {
        const Plan *p = &mb->plans[g];
        const int n = mb->nmaps[g], ngr = mb->ngr[g];
        const int mlo = f64 ? 0 : std::min(M_LO, p->mmax + 1);
        analysis_tiles(g, n, ngr, mlo, p->mmax, cm[g], target, at, acost);
        if (ngr > 1)
            for (int b = 0; b < n; ++b)
                for (int m = mlo; m <= p->mmax; ++m) rrows.push_back(MultiRow{g, m, b, 0});
        for (int b = 0; b < n; ++b)
            for (int m = 0; m <= p->mmax; ++m) prows.push_back(MultiRow{g, m, b, 0});
        if (!f64) {
            const int mt = std::min(M_LO - 1, p->mmax);
            analysis_tiles(g, n, ngb[g], 0, mt, cb[g], btarget, bt, bcost);
            for (int b = 0; b < n; ++b) {
                if (ngb[g] > 1)
                    for (int m = 0; m <= mt; ++m) brrows.push_back(MultiRow{g, m, b, 0});
                bcrows.push_back(MultiRow{g, 0, b, 0});
            }
        }
        // synthesis tiles: one m-block x 64 ring pairs x up to BM maps
        const int nw = (p->npairs + 63) / 64;
        const std::vector<int> &w0 = wst[g];
        for (int b0 = 0; b0 < n; b0 += BM) {
            const int nb = std::min(BM, n - b0);
            for (int m0 = 0; m0 <= p->mmax; m0 += MB)
                for (int w = 0; w < nw; ++w) {
                    double c = 0.0;
                    for (int m = m0; m <= std::min(p->mmax, m0 + MB - 1); ++m) {
                        const int s = std::max(w0[(size_t)m * nw + w], m);
                        if (s <= p->lmax) c += 64.0 * (p->lmax + 1 - s) * (1 + nb);
                    }
                    st.push_back(MultiTile{g, m0, MB, w, b0, nb});
                    scost.push_back(c);
                }
        }
    }
    const size_t lim = (size_t)INT32_MAX;
    if (at.size() > lim || st.size() > lim || rrows.size() > lim || prows.size() > lim || bt.size() > lim ||
        brrows.size() > lim)
        return set_error(HBX_EUNSUPPORTED, "batch too large: more than 2^31 tiles or rows");
    // launch no more workgroups than there are tiles (small batches)
    auto cap = [](int grid, size_t ntiles) {
        return (int)std::max<size_t>(NQ, std::min<size_t>((size_t)grid, (ntiles + NQ - 1) / NQ * NQ));
    };
    mb->grid_ana = cap(mb->grid_ana, at.size());
    mb->grid_syn = cap(mb->grid_syn, st.size());
    if (!f64) mb->grid_band = cap(mb->grid_band, bt.size());
    std::vector<int32_t> aq, sq, bq;
    sort_queues(at, acost, aq);
    sort_queues(st, scost, sq);
    mb->n_atiles = (int)at.size();
    mb->n_stiles = (int)st.size();
    mb->n_rrows = (int)rrows.size();
    mb->n_prows = (int)prows.size();
    if ((rc = upload_vec(mb, &mb->d_geo, geo))) return rc;
    if ((rc = upload_vec(mb, &mb->d_atiles, at))) return rc;
    if ((rc = upload_vec(mb, &mb->d_stiles, st))) return rc;
    if ((rc = upload_vec(mb, &mb->d_aq, aq))) return rc;
    if ((rc = upload_vec(mb, &mb->d_sq, sq))) return rc;
    if ((rc = upload_vec(mb, &mb->d_rrows, rrows))) return rc;
    if ((rc = upload_vec(mb, &mb->d_prows, prows))) return rc;
    if (!f64) {
        sort_queues(bt, bcost, bq);
        mb->n_btiles = (int)bt.size();
        mb->n_brrows = (int)brrows.size();
        mb->n_bcrows = (int)bcrows.size();
        if ((rc = upload_vec(mb, &mb->d_bgeo, bgeo))) return rc;
        if ((rc = upload_vec(mb, &mb->d_btiles, bt))) return rc;
        if ((rc = upload_vec(mb, &mb->d_bq, bq))) return rc;
        if ((rc = upload_vec(mb, &mb->d_brrows, brrows))) return rc;
        if ((rc = upload_vec(mb, &mb->d_bcrows, bcrows))) return rc;
    }
    return build_ring_launches(mb->plans, mb->nmaps, mb->phase_off, mb->rings, mb->band_dft, mb->ptrs);
}

static int check_call(const hbx_multi *mb, const void *const *a, const int64_t *astr, const void *const *b,
                      const int64_t *bstr, int32_t map_dtype, void *ws, uint64_t wsb, bool a_is_maps) {
    if (!mb) return set_error(HBX_EINVAL, "batch is NULL");
    if (!a || !astr || !b || !bstr) return set_error(HBX_EINVAL, "pointer or stride array is NULL");
    if (map_dtype != HBX_DT_F64 && map_dtype != HBX_DT_F32) return set_error(HBX_EINVAL, "map_dtype");
    if (!ws || wsb < mb->ws_bytes)
        return set_error(HBX_EWORKSPACE, "workspace too small: need " + std::to_string(mb->ws_bytes) + " bytes");
    int dev = -1;
    HBX_HIP(hipGetDevice(&dev));
    if (dev != mb->device)
        return set_error(HBX_EPLAN, "batch was created on device " + std::to_string(mb->device) + " but device " +
                                        std::to_string(dev) + " is current");
    for (int g = 0; g < mb->ngroups; ++g) {
        if (!a[g] || !b[g]) return set_error(HBX_EINVAL, "data pointer of group " + std::to_string(g) + " is NULL");
        const int64_t ms = a_is_maps ? astr[g] : bstr[g], as = a_is_maps ? bstr[g] : astr[g];
        if (ms < mb->plans[g].npix) return set_error(HBX_EINVAL, "map stride of group " + std::to_string(g));
        if (as < mb->plans[g].nlm) return set_error(HBX_EINVAL, "alm stride of group " + std::to_string(g));
    }
    return HBX_OK;
}

static int launch_setup(const hbx_multi *mb, void *const *alm, const int64_t *astr, void *const *maps,
                        const int64_t *mstr, char *ws, hipStream_t s) {
    MultiCall c;
    std::memset(&c, 0, sizeof(c));
    for (int g = 0; g < mb->ngroups; ++g) {
        c.alm[g] = alm[g];
        c.astride[g] = astr[g];
        c.maps[g] = maps[g];
        c.mstride[g] = mstr[g];
    }
    hipLaunchKernelGGL(multi_setup_kernel, dim3(1), dim3(64), 0, s, c, ws);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "multi_setup_kernel launch");
    return HBX_OK;
}

// per-call map pointers and strides in the workspace header
static void *const *maps_tab(char *ws) {
    return reinterpret_cast<void *const *>(ws + HDR_CALL + offsetof(MultiCall, maps));
}
static const int64_t *mstride_tab(char *ws) {
    return reinterpret_cast<const int64_t *>(ws + HDR_CALL + offsetof(MultiCall, mstride));
}

static int multi_ring_analysis(const hbx_multi *mb, int map_dtype, char *ws, hipStream_t s) {
    timer_mark(HBX_TIMER_RING_ANALYSIS, true, s);
    for (const RingLaunch &L : mb->rings) {
        const int rc = launch_multi_ring_analysis(L, mb->precision, map_dtype, maps_tab(ws), mstride_tab(ws), ws, s);
        if (rc) return rc;
    }
    timer_mark(HBX_TIMER_RING_ANALYSIS, false, s);
    return HBX_OK;
}

// FP32 batches: m < M_LO in FP64 -- band DFT of every group's rings (one
// launch), then the band's persistent Legendre launch, its reduction and the
// rounding into the call's alm
static int multi_band_analysis(const hbx_multi *mb, int map_dtype, char *ws, hipStream_t s) {
    int rc = launch_multi_ring_band(mb->band_dft, map_dtype, maps_tab(ws), mstride_tab(ws), ws, s);
    if (rc) return rc;
#define HBX_MB(RV)                                                                                          \
    hipLaunchKernelGGL((multi_analysis_kernel<double, RV>), dim3(mb->grid_band), dim3(64), 0, s, mb->d_bgeo, \
                       mb->d_btiles, mb->d_bq, ws, NQ)
    switch (mb->band_slots) {
    case 1: HBX_MB(1); break;
    case 2: HBX_MB(2); break;
    case 4: HBX_MB(4); break;
    default: return set_error(HBX_EUNSUPPORTED, "batch band slots");
    }
#undef HBX_MB
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "multi_analysis_kernel (band) launch");
    if (mb->n_brrows > 0) {
        hipLaunchKernelGGL(multi_reduce_kernel<double>, dim3(mb->n_brrows), dim3(256), 0, s, mb->d_bgeo,
                           mb->d_brrows, ws);
        if ((e = hipGetLastError()) != hipSuccess) return hip_error(e, "multi_reduce_kernel (band) launch");
    }
    hipLaunchKernelGGL(multi_band_convert_kernel, dim3(mb->n_bcrows), dim3(256), 0, s, mb->d_bgeo, mb->d_bcrows, ws);
    if ((e = hipGetLastError()) != hipSuccess) return hip_error(e, "multi_band_convert_kernel launch");
    return HBX_OK;
}

template <typename T>
static int multi_legendre_analysis(const hbx_multi *mb, char *ws, hipStream_t s) {
#define HBX_MA(RV)                                                                                        \
    hipLaunchKernelGGL((multi_analysis_kernel<T, RV>), dim3(mb->grid_ana), dim3(64), 0, s, mb->d_geo,     \
                       mb->d_atiles, mb->d_aq, ws, 0)
    timer_mark(HBX_TIMER_LEGENDRE_ANALYSIS, true, s);
    switch (mb->slots) {
    case 1: HBX_MA(1); break;
    case 2: HBX_MA(2); break;
    case 4: HBX_MA(4); break;
    case 8: HBX_MA(8); break;
    default: return set_error(HBX_EUNSUPPORTED, "batch analysis slots");
    }
#undef HBX_MA
    timer_mark(HBX_TIMER_LEGENDRE_ANALYSIS, false, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "multi_analysis_kernel launch");
    if (mb->n_rrows > 0) {
        hipLaunchKernelGGL(multi_reduce_kernel<T>, dim3(mb->n_rrows), dim3(256), 0, s, mb->d_geo, mb->d_rrows, ws);
        if ((e = hipGetLastError()) != hipSuccess) return hip_error(e, "multi_reduce_kernel launch");
    }
    return HBX_OK;
}

template <typename T>
static int multi_legendre_synthesis(const hbx_multi *mb, char *ws, hipStream_t s) {
    hipLaunchKernelGGL(multi_prescale_kernel<T>, dim3(mb->n_prows), dim3(256), 0, s, mb->d_geo, mb->d_prows, ws);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "multi_prescale_kernel launch");
#define HBX_MS(BV)                                                                                        \
    hipLaunchKernelGGL((multi_synthesis_kernel<T, BV>), dim3(mb->grid_syn), dim3(512), 0, s, mb->d_geo,   \
                       mb->d_stiles, mb->d_sq, ws)
    timer_mark(HBX_TIMER_LEGENDRE_SYNTHESIS, true, s);
    switch (mb->syn_maps) {
    case 1: HBX_MS(1); break;
    case 2: HBX_MS(2); break;
    case 4: HBX_MS(4); break;
    default: return set_error(HBX_EUNSUPPORTED, "batch synthesis maps per tile");
    }
#undef HBX_MS
    timer_mark(HBX_TIMER_LEGENDRE_SYNTHESIS, false, s);
    if ((e = hipGetLastError()) != hipSuccess) return hip_error(e, "multi_synthesis_kernel launch");
    return HBX_OK;
}

}  // namespace hbx

extern "C" {

int hbx_multi_create(const hbx_plan *const *plans, const int32_t *nmaps, int32_t ngroups,
                     const hbx_multi_tiles *tiles, hbx_multi **out) {
    if (!out) return set_error(HBX_EINVAL, "out is NULL");
    *out = nullptr;
    if (!plans || !nmaps) return set_error(HBX_EINVAL, "plans or nmaps is NULL");
    if (ngroups < 1 || ngroups > MULTI_MAXG)
        return set_error(HBX_EINVAL, "ngroups must be in [1, " + std::to_string(MULTI_MAXG) + "]");
    int dev = -1;
    HBX_HIP(hipGetDevice(&dev));
    const int prec = plans[0] ? plans[0]->impl.precision : 0;
    for (int g = 0; g < ngroups; ++g) {
        if (!plans[g]) return set_error(HBX_EINVAL, "plan of group " + std::to_string(g) + " is NULL");
        const Plan &p = plans[g]->impl;
        if (p.spin != 0) return set_error(HBX_EPLAN, "batches take spin-0 plans");
        if (p.precision != prec) return set_error(HBX_EINVAL, "every plan of a batch must have the same precision");
        if (p.device != dev)
            return set_error(HBX_EPLAN, "plan of group " + std::to_string(g) + " was created on device " +
                                            std::to_string(p.device) + " but device " + std::to_string(dev) +
                                            " is current");
        if (nmaps[g] < 1 || nmaps[g] > 65535)
            return set_error(HBX_EINVAL, "nmaps of group " + std::to_string(g) + " must be in [1, 65535]");
    }
    hbx_multi_tiles t{};
    if (tiles) t = *tiles;
    const int slots = t.valu_slots > 0 ? t.valu_slots : (prec == HBX_F64 ? 8 : 4);
    const int syn_maps = t.syn_maps > 0 ? t.syn_maps : 4;
    const int tpw = t.tiles_per_wave > 0 ? t.tiles_per_wave : 8;
    if (slots != 1 && slots != 2 && slots != 4 && slots != 8)
        return set_error(HBX_EINVAL, "valu_slots must be 1, 2, 4 or 8");
    if (syn_maps != 1 && syn_maps != 2 && syn_maps != 4) return set_error(HBX_EINVAL, "syn_maps must be 1, 2 or 4");
    if (tpw > 1024) return set_error(HBX_EINVAL, "tiles_per_wave must be <= 1024");
    auto *mb = new (std::nothrow) hbx_multi;
    if (!mb) return set_error(HBX_ENOMEM, "host allocation failed");
    mb->device = dev;
    mb->precision = prec;
    mb->ngroups = ngroups;
    mb->slots = slots;
    mb->syn_maps = syn_maps;
    mb->tiles_per_wave = tpw;
    for (int g = 0; g < ngroups; ++g) {
        mb->plans.push_back(plans[g]->impl);
        mb->nmaps.push_back(nmaps[g]);
    }
    int rc = build_multi(mb);
    if (!rc && prec == HBX_F32) {
        hipError_t e = hipStreamCreateWithFlags(&mb->side, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&mb->ev_fork, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&mb->ev_join, hipEventDisableTiming);
        if (e != hipSuccess) rc = hip_error(e, "batch side stream");
    }
    if (rc) {
        delete mb;
        return rc;
    }
    *out = mb;
    return HBX_OK;
}

int hbx_multi_destroy(hbx_multi *mb) {
    delete mb;
    return HBX_OK;
}

int hbx_multi_workspace_bytes(const hbx_multi *mb, uint64_t *bytes) {
    if (!mb || !bytes) return set_error(HBX_EINVAL, "NULL argument");
    *bytes = mb->ws_bytes;
    return HBX_OK;
}

int hbx_multi_get_info(const hbx_multi *mb, hbx_multi_info *info) {
    if (!mb || !info) return set_error(HBX_EINVAL, "NULL argument");
    std::memset(info, 0, sizeof(*info));
    info->ngroups = mb->ngroups;
    info->precision = mb->precision;
    info->valu_slots = mb->slots;
    info->syn_maps = mb->syn_maps;
    info->tiles_per_wave = mb->tiles_per_wave;
    info->analysis_tiles = mb->n_atiles;
    info->synthesis_tiles = mb->n_stiles;
    info->analysis_workgroups = mb->grid_ana;
    info->synthesis_workgroups = mb->grid_syn;
    info->workspace_bytes = mb->ws_bytes;
    return HBX_OK;
}

int hbx_multi_map2alm(const hbx_multi *mb, const void *const *maps, const int64_t *map_stride, int32_t map_dtype,
                      void *const *alm, const int64_t *alm_stride, void *ws, uint64_t ws_bytes, void *stream) {
    int rc = check_call(mb, maps, map_stride, alm, alm_stride, map_dtype, ws, ws_bytes, true);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    char *w = static_cast<char *>(ws);
    if ((rc = launch_setup(mb, alm, alm_stride, const_cast<void *const *>(maps), map_stride, w, s))) return rc;
    if (mb->precision == HBX_F64) {
        if ((rc = multi_ring_analysis(mb, map_dtype, w, s))) return rc;
        return multi_legendre_analysis<double>(mb, w, s);
    }
    // FP32: the FP64 band chain (its own buffers, the second queue counters)
    // on the side stream next to the ring FFTs and the FP32 Legendre launch;
    // forked after the setup kernel and joined before the call returns, also
    // while s is being captured (the side stream joins the capture)
    std::lock_guard<std::mutex> lock(mb->side_mu);
    HBX_HIP(hipEventRecord(mb->ev_fork, s));
    HBX_HIP(hipStreamWaitEvent(mb->side, mb->ev_fork, 0));
    const int rb = multi_band_analysis(mb, map_dtype, w, mb->side);
    HBX_HIP(hipEventRecord(mb->ev_join, mb->side));
    rc = rb ? rb : multi_ring_analysis(mb, map_dtype, w, s);
    if (!rc) rc = multi_legendre_analysis<float>(mb, w, s);
    const hipError_t ej = hipStreamWaitEvent(s, mb->ev_join, 0);
    if (rc) return rc;
    return ej == hipSuccess ? HBX_OK : hip_error(ej, "batch side stream join");
}

int hbx_multi_alm2map(const hbx_multi *mb, const void *const *alm, const int64_t *alm_stride, void *const *maps,
                      const int64_t *map_stride, int32_t map_dtype, void *ws, uint64_t ws_bytes, void *stream) {
    int rc = check_call(mb, alm, alm_stride, maps, map_stride, map_dtype, ws, ws_bytes, false);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    char *w = static_cast<char *>(ws);
    if ((rc = launch_setup(mb, const_cast<void *const *>(alm), alm_stride, maps, map_stride, w, s))) return rc;
    const bool f64 = mb->precision == HBX_F64;
    rc = f64 ? multi_legendre_synthesis<double>(mb, w, s) : multi_legendre_synthesis<float>(mb, w, s);
    if (rc) return rc;
    timer_mark(HBX_TIMER_RING_SYNTHESIS, true, s);
    for (const RingLaunch &L : mb->rings)
        if ((rc = launch_multi_ring_synthesis(L, mb->precision, map_dtype, maps_tab(w), mstride_tab(w), w, s)))
            return rc;
    timer_mark(HBX_TIMER_RING_SYNTHESIS, false, s);
    return HBX_OK;
}

}  // extern "C"
