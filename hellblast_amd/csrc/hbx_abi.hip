// C ABI of libhbx.so: argument validation, plan construction and the
// execute entry points.  Validation is host-only so it can be exercised
// without a GPU; execute functions only enqueue kernels on the given stream.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "hbx_internal.h"

namespace hbx {

static thread_local std::string g_err;

int set_error(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

int hip_error(hipError_t e, const char *what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return HBX_EDEVICE;
}

// ---------------------------------------------------------------------------
// kernel timer
// ---------------------------------------------------------------------------
constexpr int TIMER_IDS = 7;                 // HBX_TIMER_* ids are 1..6
constexpr int TIMER_MAX_LAUNCHES = 256;
struct KernelTimer {
    std::vector<hipEvent_t> ev;                 // pairs (begin, end)
    int used = 0;
    bool open = false;
};
static std::mutex g_tmu;
static uint32_t g_tmask = 0;
static int g_tdevice = -1;
static KernelTimer g_timer[TIMER_IDS];

void timer_mark(int kernel, bool begin, hipStream_t s) {
    if (kernel <= 0 || kernel >= TIMER_IDS) return;
    if (!(__atomic_load_n(&g_tmask, __ATOMIC_RELAXED) & (1u << kernel))) return;
    std::lock_guard<std::mutex> lk(g_tmu);
    if (!(g_tmask & (1u << kernel))) return;
    KernelTimer &t = g_timer[kernel];
    if (begin) {
        if (t.used + 2 > (int)t.ev.size()) return;
        t.open = hipEventRecord(t.ev[t.used], s) == hipSuccess;
    } else if (t.open) {
        if (hipEventRecord(t.ev[t.used + 1], s) == hipSuccess) t.used += 2;
        t.open = false;
    }
}

// blocks of the larger of the two layouts (2 pairs8(npairs) >= nrings)
int64_t phase_elems_per_map(const Plan *p) { return (int64_t)p->nmb * syn_block_elems(p->npairs); }
int64_t scaled_elems_per_map(const Plan *p) { return p->nrow; }

static uint64_t align256(uint64_t v) { return (v + 255) & ~uint64_t(255); }

// Phase buffer: FP32/FP64 ring coefficients for every m, plus (FP32 plans)
// the FP64 band m < M_LO.  Workspace: phase buffer | tail, where the tail
// holds either the synthesis scaled alm or the analysis partial sums (and,
// for FP32 plans, the FP64 band alm and its partial sums).
static uint64_t csize(const Plan *p) { return p->precision == HBX_F64 ? 16 : 8; }

static uint64_t phase_main_bytes(const Plan *p, int nmaps) {
    return align256((uint64_t)nmaps * phase_elems_per_map(p) * csize(p));
}

uint64_t phase_total_bytes(const Plan *p, int nmaps) {
    uint64_t b = phase_main_bytes(p, nmaps);
    if (p->precision == HBX_F32) b += align256((uint64_t)nmaps * band_phase_elems(p) * 16);
    return b;
}

void *band_phase_ptr(const Plan *p, int nmaps, const void *phase) {
    return static_cast<char *>(const_cast<void *>(phase)) + phase_main_bytes(p, nmaps);
}

uint64_t band_phase_offset(const Plan *p, int nmaps) { return phase_main_bytes(p, nmaps); }

struct AnalysisWs {
    uint64_t partial = 0, band_alm = 0, band_partial = 0, lfirst = 0, total = 0;
};

static AnalysisWs analysis_ws(const Plan *p, int nmaps) {
    AnalysisWs w;
    const uint64_t cs = csize(p);
    int g = analysis_groups(p, p->precision == HBX_F64);
    if (p->precision == HBX_F32) g = std::max(g, mfma_analysis_groups(p));
    w.partial = 0;
    uint64_t off = align256(g > 1 ? (uint64_t)nmaps * g * p->nlm * cs : 0);
    if (p->precision == HBX_F32) {
        const int g64 = std::max(band_groups(p), band_mfma_groups(p));
        w.band_alm = off;
        off += align256((uint64_t)nmaps * band_nlm(p) * 16);
        w.band_partial = off;
        off += align256(g64 > 1 ? (uint64_t)nmaps * g64 * band_nlm(p) * 16 : 0);
    }
    // first l with a nonzero partial sum, per (m, ring-pair group)
    w.lfirst = off;
    off += align256(g > 1 ? (uint64_t)(p->mmax + 1) * g * sizeof(int32_t) : 0);
    w.total = off;
    return w;
}

static uint64_t ws_tail_offset(const Plan *p, int nmaps) { return phase_total_bytes(p, nmaps); }

// Spin plans, per polarisation pair: phase of Q and U | (FP32 plans) the FP64
// band: phase, alm and partial sums of m < M_LO | tail = scaled (E, B) or the
// partial sums of both.  Synthesis puts scaled right after the phase.
struct SpinWs {
    uint64_t phase = 0, band_phase = 0, band_alm = 0, band_partial = 0, tail = 0, lfirst = 0, total = 0;
};

static SpinWs spin_ws(const Plan *p, int npol) {
    SpinWs w;
    const uint64_t cs = csize(p), n2 = 2 * (uint64_t)npol;
    const int g = spin_groups(p);
    w.phase = align256(n2 * phase_elems_per_map(p) * cs);
    uint64_t off = w.phase;
    if (p->precision == HBX_F32) {
        w.band_phase = off;
        off += align256(n2 * band_phase_elems(p) * 16);
        w.band_alm = off;
        off += align256(n2 * band_nlm(p) * 16);
        w.band_partial = off;
        off += align256(g > 1 ? n2 * g * band_nlm(p) * 16 : 0);
    }
    w.tail = off;
    const uint64_t scaled = align256(n2 * scaled_elems_per_map(p) * cs);
    // matrix-core analysis (FP32): its own ring-pair groups and the first
    // l of each group's partial sums per m
    const int gm = spin_uses_mfma(p, npol) ? spin_mfma_groups(p) : 1;
    const uint64_t part = align256(std::max(g, gm) > 1 ? n2 * std::max(g, gm) * p->nlm * cs : 0);
    w.lfirst = w.tail + part;
    const uint64_t lf = align256(gm > 1 ? (uint64_t)(p->mmax + 1) * gm * sizeof(int32_t) : 0);
    w.total = std::max(w.phase + scaled, w.lfirst + lf);
    return w;
}

static uint64_t spin_ws_bytes(const Plan *p, int npol) { return spin_ws(p, npol).total; }

static uint64_t ws_bytes(const Plan *p, int nmaps) {
    if (p->spin != 0) return spin_ws_bytes(p, nmaps);
    uint64_t scaled = align256((uint64_t)nmaps * scaled_elems_per_map(p) * csize(p));
    if (p->precision == HBX_F32) scaled = std::max(scaled, align256(syn_prep_bytes(p, nmaps)));
    const uint64_t ana = analysis_ws(p, nmaps).total;
    return ws_tail_offset(p, nmaps) + std::max(scaled, ana);
}

static int validate_geom(const hbx_geom *g, int32_t lmax, int32_t mmax) {
    if (!g) return set_error(HBX_EINVAL, "geometry is NULL");
    if (g->nrings < 1) return set_error(HBX_EINVAL, "nrings must be >= 1");
    if (!g->ring_start || !g->nphi || !g->cos_theta || !g->sin_theta)
        return set_error(HBX_EINVAL, "ring_start, nphi, cos_theta and sin_theta are required");
    if (lmax < 0) return set_error(HBX_EINVAL, "lmax must be >= 0");
    if (mmax < 0 || mmax > lmax) return set_error(HBX_EINVAL, "mmax must satisfy 0 <= mmax <= lmax");
    if (lmax > 32767) return set_error(HBX_EUNSUPPORTED, "lmax above 32767");
    const int nr = g->nrings;
    if (g->ring_start[0] < 0) return set_error(HBX_EINVAL, "ring_start[0] must be >= 0");
    for (int r = 0; r < nr; ++r) {
        const int n = g->nphi[r];
        if (n < 2 || (n & 1))
            return set_error(HBX_EINVAL, "nphi[" + std::to_string(r) + "] must be even and >= 2");
        if (g->ring_start[r + 1] - g->ring_start[r] != (int64_t)n)
            return set_error(HBX_EINVAL, "ring_start/nphi mismatch at ring " + std::to_string(r));
        const double c = g->cos_theta[r], s = g->sin_theta[r];
        if (!std::isfinite(c) || !std::isfinite(s) || s < 0.0 || std::fabs(c * c + s * s - 1.0) > 1e-10)
            return set_error(HBX_EINVAL, "cos/sin theta invalid at ring " + std::to_string(r));
        if (g->mcut && g->mcut[r] < 0) return set_error(HBX_EINVAL, "mcut must be >= 0");
        if (g->weight && !std::isfinite(g->weight[r])) return set_error(HBX_EINVAL, "weight not finite");
        if (g->phi0 && !std::isfinite(g->phi0[r])) return set_error(HBX_EINVAL, "phi0 not finite");
        const int q = nr - 1 - r;
        if (q > r) {
            if (g->nphi[q] != n)
                return set_error(HBX_EUNSUPPORTED, "rings " + std::to_string(r) + " and " + std::to_string(q) +
                                                       " are not mirror images (nphi differs)");
            if (std::fabs(g->cos_theta[q] + c) > 1e-12 || std::fabs(g->sin_theta[q] - s) > 1e-12)
                return set_error(HBX_EUNSUPPORTED, "rings " + std::to_string(r) + " and " + std::to_string(q) +
                                                       " are not mirror images (theta)");
        } else if (q == r && std::fabs(c) > 1e-12) {
            return set_error(HBX_EUNSUPPORTED, "odd middle ring must lie on the equator");
        }
    }
    return HBX_OK;
}

template <typename X>
static int upload(Plan *p, X **dst, const X *src, size_t n) {
    hipError_t e = hipMalloc(dst, n * sizeof(X));
    if (e != hipSuccess) return set_error(HBX_ENOMEM, "hipMalloc failed for plan geometry");
    p->allocations.push_back(*dst);
    p->device_bytes += n * sizeof(X);
    HBX_HIP(hipMemcpy(*dst, src, n * sizeof(X), hipMemcpyHostToDevice));
    return 0;
}

DeviceAllocs::~DeviceAllocs() {
    if (ptrs.empty() && !side) return;
    int cur = 0;
    const bool have = hipGetDevice(&cur) == hipSuccess;
    if (have && cur != device) (void)hipSetDevice(device);
    for (void *ptr : ptrs) (void)hipFree(ptr);
    if (side) (void)hipStreamDestroy(side);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (have && cur != device) (void)hipSetDevice(cur);
}

// Side stream of a plan (non-blocking, default priority) and its fork/join
// events; without them (creation failed) every call stays on its own stream.
static void create_side_stream(DeviceAllocs *d) {
    if (hipStreamCreateWithFlags(&d->side, hipStreamNonBlocking) != hipSuccess) d->side = nullptr;
    if (d->side && (hipEventCreateWithFlags(&d->ev_fork, hipEventDisableTiming) != hipSuccess ||
                    hipEventCreateWithFlags(&d->ev_join, hipEventDisableTiming) != hipSuccess)) {
        (void)hipStreamDestroy(d->side);
        d->side = nullptr;
    }
    (void)hipGetLastError();
}

// Independent work of one call (the FP64 band of FP32 plans) on the plan's
// side stream, concurrent with the main work on the caller's stream s: the
// constructor makes the side stream wait for the work already queued on s,
// join() (or the destructor) makes s wait for the side work.  The plan's
// lock is held from fork to join because the two events are shared; while
// s is being captured into a graph (or there is no side stream, or use is
// false) side() is s and nothing forks.
class SideFork {
  public:
    SideFork(const Plan *p, hipStream_t s, bool use = true) : d_(p->allocations.s.get()), s_(s), side_(s) {
        if (!d_->side || !use) return;
        hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &st) != hipSuccess || st != hipStreamCaptureStatusNone) {
            (void)hipGetLastError();
            return;
        }
        lock_ = std::unique_lock<std::mutex>(d_->side_mu);
        if (hipEventRecord(d_->ev_fork, s) != hipSuccess || hipStreamWaitEvent(d_->side, d_->ev_fork, 0) != hipSuccess) {
            (void)hipGetLastError();
            lock_.unlock();
            return;
        }
        side_ = d_->side;
    }
    SideFork(const SideFork &) = delete;
    SideFork &operator=(const SideFork &) = delete;
    ~SideFork() { (void)join(); }
    hipStream_t side() const { return side_; }
    int join() {
        if (side_ == s_) return HBX_OK;
        hipError_t e = hipEventRecord(d_->ev_join, side_);
        if (e == hipSuccess) e = hipStreamWaitEvent(s_, d_->ev_join, 0);
        side_ = s_;
        lock_.unlock();
        return e == hipSuccess ? HBX_OK : hip_error(e, "side stream join");
    }

  private:
    DeviceAllocs *d_;
    hipStream_t s_, side_;
    std::unique_lock<std::mutex> lock_;
};

// drops this plan's reference to the shared device tables
static void destroy_plan(Plan *p) { p->allocations.s.reset(); }

static int build_plan(Plan *p, const hbx_geom *g) {
    const int nr = p->nrings;
    std::vector<int32_t> nphi(g->nphi, g->nphi + nr);
    std::vector<int32_t> mcut(nr);
    std::vector<double> scale(nr);
    bool any_phi0 = false;
    for (int r = 0; r < nr; ++r) {
        mcut[r] = g->mcut ? g->mcut[r] : g->nphi[r];
        scale[r] = (g->weight ? g->weight[r] : 1.0) / (double)g->nphi[r];
        if (g->phi0 && g->phi0[r] != 0.0) any_phi0 = true;
    }
    int rc;
    if ((rc = upload(p, &p->d_ring_start, g->ring_start, nr + 1))) return rc;
    if ((rc = upload(p, &p->d_nphi, nphi.data(), nr))) return rc;
    if ((rc = upload(p, &p->d_mcut, mcut.data(), nr))) return rc;
    if ((rc = upload(p, &p->d_scale_ana, scale.data(), nr))) return rc;
    p->d_scale_adj = p->d_scale_ana;
    if (any_phi0 && (rc = upload(p, &p->d_phi0, g->phi0, nr))) return rc;
    const int np = p->npairs;
    std::vector<int32_t> pn(np), ps(np);
    std::vector<double> px(np), psin(np);
    for (int i = 0; i < np; ++i) {
        pn[i] = i;
        ps[i] = nr - 1 - i;
        px[i] = g->cos_theta[i];
        psin[i] = g->sin_theta[i];
    }
    if ((rc = upload(p, &p->d_pair_n, pn.data(), np))) return rc;
    if ((rc = upload(p, &p->d_pair_s, ps.data(), np))) return rc;
    if ((rc = upload(p, &p->d_px, px.data(), np))) return rc;
    // FP32 spin plans analyse Q/U rings in FP64 (output rounded once to FP32)
    p->ring_f64 = p->precision == HBX_F32 && p->spin != 0;
    if ((rc = plan_build_fft_tables(p, nphi, any_phi0 ? g->phi0 : nullptr))) return rc;
    // switch-on tolerance: FP64 plans 2^-64, FP32 plans 2^-40 (relative
    // contributions far below each format's rounding error)
    const double tol = p->precision == HBX_F64 ? -64.0 : -40.0;
    if ((rc = plan_build_recursion(p, px, psin, tol))) return rc;
    if (p->precision == HBX_F32 && p->spin == 0 && (rc = prepare_mfma_kernels())) return rc;
    if (p->precision == HBX_F32 && p->spin == 0 && (rc = plan_build_syn_chunks(p))) return rc;
    return 0;
}

static int check_exec(const hbx_plan *plan, int32_t nmaps, const void *a, const void *b, void *ws,
                      uint64_t wsb, int need_ws, bool spin_entry = false) {
    if (!plan) return set_error(HBX_EINVAL, "plan is NULL");
    if (spin_entry && plan->impl.spin == 0)
        return set_error(HBX_EPLAN, "spin entry point called with a spin-0 plan");
    if (!spin_entry && plan->impl.spin != 0)
        return set_error(HBX_EPLAN, "spin-weighted plan: use the hbx_*_spin entry points");
    const int nmax = spin_entry ? 32767 : 65535;
    if (nmaps < 0 || nmaps > nmax)
        return set_error(HBX_EINVAL, "nmaps must be in [0, " + std::to_string(nmax) + "]");
    if (nmaps == 0) return HBX_OK;
    if (!a || !b) return set_error(HBX_EINVAL, "data pointer is NULL");
    if (need_ws) {
        const uint64_t need = ws_bytes(&plan->impl, nmaps);
        if (!ws || wsb < need)
            return set_error(HBX_EWORKSPACE, "workspace too small: need " + std::to_string(need) + " bytes");
    }
    int dev = -1;
    HBX_HIP(hipGetDevice(&dev));
    if (dev != plan->impl.device)
        return set_error(HBX_EPLAN, "plan was created on device " + std::to_string(plan->impl.device) +
                                        " but device " + std::to_string(dev) + " is current");
    return HBX_OK;
}

// Ring stage of an analysis: FFTs for every m, plus the FP64 band of FP32 plans
int ring_analysis(const Plan *p, int nmaps, const void *maps, int map_dtype, int64_t map_stride, void *phase,
                         const double *scale, hipStream_t s) {
    if (p->precision != HBX_F32) return launch_ring_analysis(p, nmaps, maps, map_dtype, map_stride, phase, scale, s);
    // the band reads the same maps and writes its own buffer: side stream
    SideFork fork(p, s);
    int rc = launch_ring_band(p, nmaps, maps, map_dtype, map_stride, band_phase_ptr(p, nmaps, phase), scale,
                              fork.side());
    if (!rc) rc = launch_ring_analysis(p, nmaps, maps, map_dtype, map_stride, phase, scale, s);
    const int rj = fork.join();
    return rc ? rc : rj;
}

// Legendre analysis: matrix cores for FP32 batches, VALU otherwise; FP32
// plans take m < M_LO from the FP64 band.  tail: workspace after the phase.
static int legendre_analysis(const Plan *p, int nmaps, const void *phase, void *alm, int64_t alm_stride,
                             void *tail, int adjoint, hipStream_t s) {
    const AnalysisWs w = analysis_ws(p, nmaps);
    char *t = static_cast<char *>(tail);
    if (p->precision != HBX_F32)
        return launch_leg_analysis(p, nmaps, phase, alm, alm_stride, t + w.partial, adjoint, s);
    // the band (alm rows m < M_LO, own workspace) on the side stream; the
    // main kernels write rows m >= M_LO
    SideFork fork(p, s);
    int rc = launch_leg_analysis_band(p, nmaps, band_phase_ptr(p, nmaps, phase), alm, alm_stride, t + w.band_alm,
                                      t + w.band_partial, adjoint, fork.side());
    if (!rc) {
        if (nmaps >= mfma_min_maps(p))
            rc = launch_leg_analysis_mfma(p, nmaps, phase, alm, alm_stride, t + w.partial,
                                          reinterpret_cast<int32_t *>(t + w.lfirst), adjoint, s);
        else
            rc = launch_leg_analysis(p, nmaps, phase, alm, alm_stride, t + w.partial, adjoint, s);
    }
    const int rj = fork.join();
    return rc ? rc : rj;
}

// Legendre synthesis: matrix cores for FP32 batches, VALU otherwise
static int legendre_synthesis(const Plan *p, int nmaps, const void *alm, int64_t alm_stride, void *phase,
                              void *scaled, int adjoint, hipStream_t s) {
    if (p->precision == HBX_F32 && nmaps >= mfma_min_maps(p))
        return launch_leg_synthesis_mfma(p, nmaps, alm, alm_stride, phase, scaled, adjoint, s);
    int rc = launch_prescale(p, nmaps, alm, alm_stride, scaled, adjoint, s);
    if (rc) return rc;
    return launch_leg_synthesis(p, nmaps, scaled, phase, s);
}

static int check_strides(const Plan *p, int32_t map_dtype, int64_t map_stride, int64_t alm_stride) {
    if (map_dtype != HBX_DT_F64 && map_dtype != HBX_DT_F32) return set_error(HBX_EINVAL, "map_dtype");
    if (map_stride < p->npix) return set_error(HBX_EINVAL, "map_stride smaller than npix");
    if (alm_stride < p->nlm) return set_error(HBX_EINVAL, "alm_stride smaller than the alm count");
    return HBX_OK;
}

}  // namespace hbx

using namespace hbx;

extern "C" {

int hbx_abi_version(void) { return HBX_ABI_VERSION; }

const char *hbx_last_error(void) { return g_err.c_str(); }

int hbx_validate_geom(const hbx_geom *g, int32_t lmax, int32_t mmax) { return validate_geom(g, lmax, mmax); }

int hbx_plan_create(const hbx_geom *g, int32_t lmax, int32_t mmax, int32_t spin, int32_t precision,
                    uint32_t flags, hbx_plan **out) {
    if (!out) return set_error(HBX_EINVAL, "out is NULL");
    *out = nullptr;
    int rc = validate_geom(g, lmax, mmax);
    if (rc) return rc;
    if (precision != HBX_F64 && precision != HBX_F32) return set_error(HBX_EINVAL, "precision must be 0 or 1");
    if (spin < 0 || spin > HBX_MAX_SPIN)
        return set_error(HBX_EINVAL, "spin must be in [0, " + std::to_string(HBX_MAX_SPIN) + "]");
    if (spin > lmax) return set_error(HBX_EINVAL, "spin must not exceed lmax");
    if (flags != 0) return set_error(HBX_EINVAL, "flags must be 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return set_error(HBX_EDEVICE, "no HIP device");
    auto *hp = new hbx_plan();
    Plan *p = &hp->impl;
    if (hipGetDevice(&p->device) != hipSuccess) {
        delete hp;
        return set_error(HBX_EDEVICE, "hipGetDevice failed");
    }
    p->allocations.s->device = p->device;
    p->nrings = g->nrings;
    p->npairs = (g->nrings + 1) / 2;
    p->lmax = lmax;
    p->mmax = mmax;
    p->nmb = (mmax + 1 + MB - 1) / MB;
    p->spin = spin;
    p->precision = precision;
    p->npix = g->ring_start[g->nrings];
    p->nlm = nalm(lmax, mmax);
    rc = build_plan(p, g);
    if (rc) {
        const std::string msg = g_err;
        destroy_plan(p);
        delete hp;
        return set_error(rc, msg);
    }
    if (p->precision == HBX_F32) create_side_stream(p->allocations.s.get());
    *out = hp;
    return HBX_OK;
}

int hbx_plan_destroy(hbx_plan *plan) {
    if (!plan) return HBX_OK;
    destroy_plan(&plan->impl);
    delete plan;
    return HBX_OK;
}

int hbx_plan_get_info(const hbx_plan *plan, hbx_plan_info *info) {
    if (!plan || !info) return set_error(HBX_EINVAL, "NULL argument");
    const Plan &p = plan->impl;
    info->nrings = p.nrings;
    info->npairs = p.npairs;
    info->lmax = p.lmax;
    info->mmax = p.mmax;
    info->precision = p.precision;
    info->spin = p.spin;
    info->npix = p.npix;
    info->nalm = p.nlm;
    info->device_bytes = p.device_bytes;
    info->fp64_band_m = (p.precision == HBX_F32 && p.spin == 0) ? std::min(M_LO, p.mmax + 1) : 0;
    info->reserved = 0;
    return HBX_OK;
}

int hbx_plan_derive(const hbx_plan *base, const hbx_tiles *t, hbx_plan **out) {
    if (!out) return set_error(HBX_EINVAL, "out is NULL");
    *out = nullptr;
    if (!base || !t) return set_error(HBX_EINVAL, "NULL argument");
    auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
    if (t->valu_slots != 0 && (!pow2(t->valu_slots) || t->valu_slots > 16))
        return set_error(HBX_EINVAL, "valu_slots must be 0 or a power of two <= 16");
    if (t->mfma_waves != 0 && t->mfma_waves != 4 && t->mfma_waves != 8 && t->mfma_waves != 12 &&
        t->mfma_waves != 16)
        return set_error(HBX_EINVAL, "mfma_waves must be 0, 4, 8, 12 or 16");
    if (t->mfma_min_maps < 0) return set_error(HBX_EINVAL, "mfma_min_maps must be >= 0");
    if (t->spin_slots != 0 && t->spin_slots != 1 && t->spin_slots != 2 && t->spin_slots != 4)
        return set_error(HBX_EINVAL, "spin_slots must be 0, 1, 2 or 4");
    if (t->mfma_kind != 0 && t->mfma_kind != HBX_MFMA_F32 && t->mfma_kind != HBX_MFMA_F16X2)
        return set_error(HBX_EINVAL, "mfma_kind must be 0, HBX_MFMA_F32 or HBX_MFMA_F16X2");
    if (t->mfma_waves == 12 && t->mfma_kind != HBX_MFMA_F32)
        return set_error(HBX_EINVAL, "mfma_waves 12 needs mfma_kind HBX_MFMA_F32");
    if (t->mfma_groups < 0 || t->mfma_groups > 2) return set_error(HBX_EINVAL, "mfma_groups must be 0, 1 or 2");
    if (t->mfma_waves == 16 && t->mfma_kind == HBX_MFMA_F32)
        return set_error(HBX_EINVAL, "mfma_waves 16 needs the split-fp16 matrix-core kernels");
    auto *hp = new hbx_plan();
    hp->impl = base->impl;            // shares the device tables (reference counted)
    Plan &p = hp->impl;
    p.tile_valu_slots = t->valu_slots;
    p.tile_mfma_waves = t->mfma_waves;
    p.tile_mfma_min_maps = t->mfma_min_maps;
    p.tile_spin_slots = t->spin_slots;
    p.tile_mfma_kind = t->mfma_kind;
    p.tile_syn_groups = t->mfma_groups;
    *out = hp;
    return HBX_OK;
}

int hbx_plan_get_tiles(const hbx_plan *plan, hbx_tiles *t) {
    if (!plan || !t) return set_error(HBX_EINVAL, "NULL argument");
    const Plan &p = plan->impl;
    t->valu_slots = analysis_slots_of(&p, p.precision == HBX_F64);
    t->mfma_waves = mfma_analysis_waves(&p);
    t->mfma_min_maps = mfma_min_maps(&p);
    t->spin_slots = spin_slots_of(&p);
    t->mfma_kind = mfma_kind_of(&p);
    t->mfma_groups = p.tile_syn_groups;       // 0: chosen per batch
    return HBX_OK;
}

int hbx_workspace_bytes(const hbx_plan *plan, int32_t nmaps, uint64_t *bytes) {
    if (!plan || !bytes) return set_error(HBX_EINVAL, "NULL argument");
    if (nmaps < 0) return set_error(HBX_EINVAL, "nmaps must be >= 0");
    *bytes = ws_bytes(&plan->impl, nmaps);
    return HBX_OK;
}

// Whole analysis (map2alm and its adjoint).  FP32 plans run the FP64 band
// chain -- band DFT, band Legendre, rounding into alm rows m < M_LO -- on the
// side stream beside the ring FFTs, whose barrier-bound workgroups leave
// issue slots free; forked at the Legendre stage it would sit beside the
// split-fp16 analysis, which fills the GPU, for that kernel's whole run.
static int analysis(const Plan *p, int nmaps, const void *maps, int map_dtype, int64_t map_stride, void *alm,
                    int64_t alm_stride, void *ws, const double *scale, int adjoint, hipStream_t s) {
    void *tail = static_cast<char *>(ws) + ws_tail_offset(p, nmaps);
    if (p->precision != HBX_F32) {
        int rc = launch_ring_analysis(p, nmaps, maps, map_dtype, map_stride, ws, scale, s);
        if (rc) return rc;
        return legendre_analysis(p, nmaps, ws, alm, alm_stride, tail, adjoint, s);
    }
    const AnalysisWs w = analysis_ws(p, nmaps);
    char *t = static_cast<char *>(tail);
    void *band = band_phase_ptr(p, nmaps, ws);
    // below 2^17 pixels per call the fork and join cost more than the
    // overlap gains (one map, FP32 map2alm, with / without the fork: nside 64
    // 0.075 / 0.050 ms, nside 128 0.080 / 0.087 ms)
    SideFork fork(p, s, p->npix * nmaps >= ((int64_t)1 << 17));
    int rc = launch_ring_band(p, nmaps, maps, map_dtype, map_stride, band, scale, fork.side());
    if (!rc)
        rc = launch_leg_analysis_band(p, nmaps, band, alm, alm_stride, t + w.band_alm, t + w.band_partial, adjoint,
                                      fork.side());
    if (!rc) rc = launch_ring_analysis(p, nmaps, maps, map_dtype, map_stride, ws, scale, s);
    if (!rc) {
        if (nmaps >= mfma_min_maps(p))
            rc = launch_leg_analysis_mfma(p, nmaps, ws, alm, alm_stride, t + w.partial,
                                          reinterpret_cast<int32_t *>(t + w.lfirst), adjoint, s);
        else
            rc = launch_leg_analysis(p, nmaps, ws, alm, alm_stride, t + w.partial, adjoint, s);
    }
    const int rj = fork.join();
    return rc ? rc : rj;
}

int hbx_map2alm(const hbx_plan *plan, int32_t nmaps, const void *maps, int32_t map_dtype, int64_t map_stride,
                void *alm, int64_t alm_stride, void *ws, uint64_t ws_bytes_, void *stream) {
    int rc = check_exec(plan, nmaps, maps, alm, ws, ws_bytes_, 1);
    if (rc || nmaps == 0) return rc;
    const Plan *p = &plan->impl;
    if ((rc = check_strides(p, map_dtype, map_stride, alm_stride))) return rc;
    return analysis(p, nmaps, maps, map_dtype, map_stride, alm, alm_stride, ws, p->d_scale_ana, 0,
                    (hipStream_t)stream);
}

int hbx_alm2map_adjoint(const hbx_plan *plan, int32_t nmaps, const void *maps, int32_t map_dtype,
                        int64_t map_stride, void *alm, int64_t alm_stride, void *ws, uint64_t ws_bytes_,
                        void *stream) {
    int rc = check_exec(plan, nmaps, maps, alm, ws, ws_bytes_, 1);
    if (rc || nmaps == 0) return rc;
    const Plan *p = &plan->impl;
    if ((rc = check_strides(p, map_dtype, map_stride, alm_stride))) return rc;
    return analysis(p, nmaps, maps, map_dtype, map_stride, alm, alm_stride, ws, nullptr, 1, (hipStream_t)stream);
}

static int synth(const hbx_plan *plan, int32_t nmaps, const void *alm, int64_t alm_stride, void *maps,
                 int32_t map_dtype, int64_t map_stride, void *ws, uint64_t ws_bytes_, void *stream,
                 int adjoint) {
    int rc = check_exec(plan, nmaps, alm, maps, ws, ws_bytes_, 1);
    if (rc || nmaps == 0) return rc;
    const Plan *p = &plan->impl;
    if ((rc = check_strides(p, map_dtype, map_stride, alm_stride))) return rc;
    hipStream_t s = (hipStream_t)stream;
    void *phase = ws;
    void *scaled = static_cast<char *>(ws) + ws_tail_offset(p, nmaps);
    if ((rc = legendre_synthesis(p, nmaps, alm, alm_stride, phase, scaled, adjoint, s))) return rc;
    return launch_ring_synthesis(p, nmaps, phase, maps, map_dtype, map_stride, adjoint ? p->d_scale_adj : nullptr, s);
}

int hbx_alm2map(const hbx_plan *plan, int32_t nmaps, const void *alm, int64_t alm_stride, void *maps,
                int32_t map_dtype, int64_t map_stride, void *ws, uint64_t ws_bytes_, void *stream) {
    return synth(plan, nmaps, alm, alm_stride, maps, map_dtype, map_stride, ws, ws_bytes_, stream, 0);
}

int hbx_map2alm_adjoint(const hbx_plan *plan, int32_t nmaps, const void *alm, int64_t alm_stride, void *maps,
                        int32_t map_dtype, int64_t map_stride, void *ws, uint64_t ws_bytes_, void *stream) {
    return synth(plan, nmaps, alm, alm_stride, maps, map_dtype, map_stride, ws, ws_bytes_, stream, 1);
}

// ---------------------------------------------------------------------------
// spin entry points: maps 2b / 2b+1 are Q / U and alm 2b / 2b+1 are E / B of
// polarisation pair b, so npol pairs take 2*npol maps and 2*npol alm rows
// ---------------------------------------------------------------------------
static int spin_analysis(const hbx_plan *plan, int32_t npol, const void *maps, int32_t map_dtype, int64_t map_stride,
                         void *alm, int64_t alm_stride, void *ws, uint64_t ws_bytes_, void *stream, int adjoint) {
    int rc = check_exec(plan, npol, maps, alm, ws, ws_bytes_, 1, true);
    if (rc || npol == 0) return rc;
    const Plan *p = &plan->impl;
    if ((rc = check_strides(p, map_dtype, map_stride, alm_stride))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const SpinWs w = spin_ws(p, npol);
    char *base = static_cast<char *>(ws);
    const double *scale = adjoint ? nullptr : p->d_scale_ana;
    const bool mf = spin_uses_mfma(p, npol);
    auto main_work = [&]() {
        int r = launch_ring_analysis(p, 2 * npol, maps, map_dtype, map_stride, ws, scale, s);
        if (r) return r;
        if (mf)
            return launch_spin_analysis_mfma(p, npol, ws, alm, alm_stride, base + w.tail,
                                             reinterpret_cast<int32_t *>(base + w.lfirst), adjoint, s);
        return launch_spin_analysis(p, npol, ws, alm, alm_stride, base + w.tail, adjoint, s);
    };
    // FP32 plans: m < M_LO in FP64 (the low-m rows carry most of the power of
    // red spectra): ring DFT of the maps, then E/B rows m < M_LO from the
    // band's own workspace
    auto band_work = [&](hipStream_t bs) {
        int r = launch_ring_band(p, 2 * npol, maps, map_dtype, map_stride, base + w.band_phase, scale, bs);
        if (r) return r;
        return launch_spin_analysis_band(p, npol, base + w.band_phase, alm, alm_stride, base + w.band_alm,
                                         base + w.band_partial, adjoint, bs);
    };
    if (p->precision != HBX_F32) return main_work();
    if (!mf) {
        // the VALU kernels write every row m, the band's rows then replace m < M_LO
        if ((rc = main_work())) return rc;
        return band_work(s);
    }
    // the matrix-core kernels write rows m >= M_LO only: the band runs whole
    // on the side stream
    SideFork fork(p, s);
    rc = band_work(fork.side());
    if (!rc) rc = main_work();
    const int rj = fork.join();
    return rc ? rc : rj;
}

static int spin_synthesis(const hbx_plan *plan, int32_t npol, const void *alm, int64_t alm_stride, void *maps,
                          int32_t map_dtype, int64_t map_stride, void *ws, uint64_t ws_bytes_, void *stream,
                          int adjoint) {
    int rc = check_exec(plan, npol, alm, maps, ws, ws_bytes_, 1, true);
    if (rc || npol == 0) return rc;
    const Plan *p = &plan->impl;
    if ((rc = check_strides(p, map_dtype, map_stride, alm_stride))) return rc;
    hipStream_t s = (hipStream_t)stream;
    const uint64_t phase = align256(2 * (uint64_t)npol * phase_elems_per_map(p) * csize(p));
    void *scaled = static_cast<char *>(ws) + phase;
    if ((rc = launch_spin_prescale(p, npol, alm, alm_stride, scaled, adjoint, s))) return rc;
    if ((rc = launch_spin_synthesis(p, npol, scaled, ws, s))) return rc;
    return launch_ring_synthesis(p, 2 * npol, ws, maps, map_dtype, map_stride, adjoint ? p->d_scale_adj : nullptr, s);
}

int hbx_map2alm_spin(const hbx_plan *plan, int32_t npol, const void *maps, int32_t map_dtype, int64_t map_stride,
                     void *alm, int64_t alm_stride, void *ws, uint64_t ws_bytes_, void *stream) {
    return spin_analysis(plan, npol, maps, map_dtype, map_stride, alm, alm_stride, ws, ws_bytes_, stream, 0);
}

int hbx_alm2map_spin_adjoint(const hbx_plan *plan, int32_t npol, const void *maps, int32_t map_dtype,
                             int64_t map_stride, void *alm, int64_t alm_stride, void *ws, uint64_t ws_bytes_,
                             void *stream) {
    return spin_analysis(plan, npol, maps, map_dtype, map_stride, alm, alm_stride, ws, ws_bytes_, stream, 1);
}

int hbx_alm2map_spin(const hbx_plan *plan, int32_t npol, const void *alm, int64_t alm_stride, void *maps,
                     int32_t map_dtype, int64_t map_stride, void *ws, uint64_t ws_bytes_, void *stream) {
    return spin_synthesis(plan, npol, alm, alm_stride, maps, map_dtype, map_stride, ws, ws_bytes_, stream, 0);
}

int hbx_map2alm_spin_adjoint(const hbx_plan *plan, int32_t npol, const void *alm, int64_t alm_stride, void *maps,
                             int32_t map_dtype, int64_t map_stride, void *ws, uint64_t ws_bytes_, void *stream) {
    return spin_synthesis(plan, npol, alm, alm_stride, maps, map_dtype, map_stride, ws, ws_bytes_, stream, 1);
}

static uint64_t phase_bytes_of(const Plan *p, int nmaps) { return phase_total_bytes(p, nmaps); }

int hbx_phase_bytes(const hbx_plan *plan, int32_t nmaps, uint64_t *bytes) {
    if (!plan || !bytes) return set_error(HBX_EINVAL, "NULL argument");
    if (nmaps < 0) return set_error(HBX_EINVAL, "nmaps must be >= 0");
    *bytes = phase_bytes_of(&plan->impl, nmaps);
    return HBX_OK;
}

static int check_phase(const Plan *p, int nmaps, const void *phase, uint64_t phase_bytes) {
    if (!phase || phase_bytes < phase_bytes_of(p, nmaps))
        return set_error(HBX_EWORKSPACE, "phase buffer too small: need " + std::to_string(phase_bytes_of(p, nmaps)) +
                                             " bytes");
    return HBX_OK;
}

int hbx_ring_analysis(const hbx_plan *plan, int32_t nmaps, const void *maps, int32_t map_dtype, int64_t map_stride,
                      void *phase, uint64_t phase_bytes, void *stream) {
    int rc = check_exec(plan, nmaps, maps, phase, nullptr, 0, 0);
    if (rc || nmaps == 0) return rc;
    const Plan *p = &plan->impl;
    if ((rc = check_strides(p, map_dtype, map_stride, p->nlm))) return rc;
    if ((rc = check_phase(p, nmaps, phase, phase_bytes))) return rc;
    return ring_analysis(p, nmaps, maps, map_dtype, map_stride, phase, p->d_scale_ana, (hipStream_t)stream);
}

int hbx_legendre_analysis(const hbx_plan *plan, int32_t nmaps, const void *phase, uint64_t phase_bytes, void *alm,
                          int64_t alm_stride, void *ws, uint64_t ws_bytes_, void *stream) {
    int rc = check_exec(plan, nmaps, phase, alm, ws, ws_bytes_, 1);
    if (rc || nmaps == 0) return rc;
    const Plan *p = &plan->impl;
    if ((rc = check_strides(p, HBX_DT_F64, p->npix, alm_stride))) return rc;
    if ((rc = check_phase(p, nmaps, phase, phase_bytes))) return rc;
    return legendre_analysis(p, nmaps, phase, alm, alm_stride, ws, 0, (hipStream_t)stream);
}

int hbx_legendre_synthesis(const hbx_plan *plan, int32_t nmaps, const void *alm, int64_t alm_stride, void *phase,
                           uint64_t phase_bytes, void *ws, uint64_t ws_bytes_, void *stream) {
    int rc = check_exec(plan, nmaps, alm, phase, ws, ws_bytes_, 1);
    if (rc || nmaps == 0) return rc;
    const Plan *p = &plan->impl;
    if ((rc = check_strides(p, HBX_DT_F64, p->npix, alm_stride))) return rc;
    if ((rc = check_phase(p, nmaps, phase, phase_bytes))) return rc;
    return legendre_synthesis(p, nmaps, alm, alm_stride, phase, ws, 0, (hipStream_t)stream);
}

int hbx_ring_synthesis(const hbx_plan *plan, int32_t nmaps, const void *phase, uint64_t phase_bytes, void *maps,
                       int32_t map_dtype, int64_t map_stride, void *stream) {
    int rc = check_exec(plan, nmaps, phase, maps, nullptr, 0, 0);
    if (rc || nmaps == 0) return rc;
    const Plan *p = &plan->impl;
    if ((rc = check_strides(p, map_dtype, map_stride, p->nlm))) return rc;
    if ((rc = check_phase(p, nmaps, phase, phase_bytes))) return rc;
    return launch_ring_synthesis(p, nmaps, phase, maps, map_dtype, map_stride, nullptr, (hipStream_t)stream);
}

int hbx_timer_arm(uint32_t mask) {
    if (mask & ~(((1u << TIMER_IDS) - 1) & ~1u)) return set_error(HBX_EINVAL, "timer mask has unknown bits");
    std::lock_guard<std::mutex> lk(g_tmu);
    __atomic_store_n(&g_tmask, 0u, __ATOMIC_RELAXED);
    if (mask == 0) return HBX_OK;
    int dev = -1;
    HBX_HIP(hipGetDevice(&dev));
    if (dev != g_tdevice) {
        for (KernelTimer &t : g_timer) {
            for (hipEvent_t e : t.ev) (void)hipEventDestroy(e);
            t.ev.clear();
        }
        g_tdevice = dev;
    }
    for (int k = 1; k < TIMER_IDS; ++k) {
        if (!(mask & (1u << k))) continue;
        KernelTimer &t = g_timer[k];
        while ((int)t.ev.size() < 2 * TIMER_MAX_LAUNCHES) {
            hipEvent_t e;
            HBX_HIP(hipEventCreate(&e));
            t.ev.push_back(e);
        }
        t.used = 0;
        t.open = false;
    }
    __atomic_store_n(&g_tmask, mask, __ATOMIC_RELAXED);
    return HBX_OK;
}

int hbx_timer_read(int32_t kernel, int32_t *launches, double *total_ms) {
    if (!launches || !total_ms) return set_error(HBX_EINVAL, "NULL argument");
    if (kernel <= 0 || kernel >= TIMER_IDS) return set_error(HBX_EINVAL, "timer kernel id");
    std::lock_guard<std::mutex> lk(g_tmu);
    KernelTimer &t = g_timer[kernel];
    double sum = 0.0;
    for (int i = 0; i < t.used; i += 2) {
        HBX_HIP(hipEventSynchronize(t.ev[i + 1]));
        float ms = 0.0f;
        HBX_HIP(hipEventElapsedTime(&ms, t.ev[i], t.ev[i + 1]));
        sum += ms;
    }
    *launches = t.used / 2;
    *total_ms = sum;
    t.used = 0;
    t.open = false;
    return HBX_OK;
}

int hbx_alm2cl(int32_t lmax, int32_t mmax, int32_t nmaps, const void *alm, int32_t alm_dtype, int64_t alm_stride,
               double *cl, int64_t cl_stride, void *stream) {
    if (lmax < 0 || mmax < 0 || mmax > lmax) return set_error(HBX_EINVAL, "need 0 <= mmax <= lmax");
    if (nmaps < 0 || nmaps > 65535) return set_error(HBX_EINVAL, "nmaps must be in [0, 65535]");
    if (nmaps == 0) return HBX_OK;
    if (!alm || !cl) return set_error(HBX_EINVAL, "data pointer is NULL");
    if (alm_dtype != HBX_DT_F64 && alm_dtype != HBX_DT_F32) return set_error(HBX_EINVAL, "alm_dtype");
    if (alm_stride < nalm(lmax, mmax) || cl_stride < lmax + 1) return set_error(HBX_EINVAL, "stride too small");
    return launch_alm2cl(lmax, mmax, nmaps, alm, alm_dtype, alm_stride, cl, cl_stride, (hipStream_t)stream);
}

}  // extern "C"
