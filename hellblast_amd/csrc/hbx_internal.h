// Internal declarations shared by the libhbx.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hbx.h"

namespace hbx {

// ---------------------------------------------------------------------------
// error reporting: thread-local message + status code
// ---------------------------------------------------------------------------
int set_error(int code, const std::string &msg);
int hip_error(hipError_t e, const char *what);

#define HBX_HIP(call)                                                        \
    do {                                                                     \
        hipError_t _e = (call);                                              \
        if (_e != hipSuccess) return ::hbx::hip_error(_e, #call);            \
    } while (0)

// kernel timer: brackets the launches of the armed kernel id (HBX_TIMER_*)
// with HIP events on their stream; a no-op unless that id is armed
void timer_mark(int kernel, bool begin, hipStream_t s);

// ---------------------------------------------------------------------------
// sizes
// ---------------------------------------------------------------------------
constexpr int MB = 8;          // m values per m-block of the phase buffer
constexpr int CHUNK = 8;       // l steps per recursion chunk
constexpr int TAILPAD = 128;   // zero coefficients after the last recursion row
// FP32 plans analyse m < M_LO in FP64: the ring coefficients of low m carry
// the large-scale power, and rounding them to FP32 costs ~1e-5 of sqrt(C_l)
// at high l for red spectra (measured at nside 2048, lmax 4095, C_l =
// 1/(l+1)^2: eps_max 1.45e-5 with no band, 3.5e-6 with m < 8 or m < 16).
constexpr int M_LO = 8;
constexpr int NMB_LO = M_LO / MB;

// Phase buffer layouts.  Analysis direction (ring analysis writes, Legendre
// analysis reads): [map][m/8][ring][m%8], m-blocks nr * MB elements apart;
// one ring workgroup writes 8 m of its ring per 64-byte line.  Synthesis
// direction (Legendre synthesis writes, ring synthesis reads):
// [map][m/8][half][pair/8][m%8][pair%8], half 0 holding ring p and half 1
// ring nr-1-p of ring pair p (half 1 has no entry for the ring pair of one
// ring); one Legendre workgroup (one m, 64 or more ring pairs) writes 8 ring
// pairs of one m per 64-byte line.  Maps lie phase_elems_per_map() apart in
// both directions.
__host__ __device__ inline int pairs8(int npairs) { return (npairs + 7) & ~7; }
__host__ __device__ inline int64_t syn_block_elems(int npairs) { return (int64_t)2 * pairs8(npairs) * MB; }
// offset of ring pair p, half (0: ring p, 1: ring nr-1-p), m in one map
__host__ __device__ inline int64_t syn_ofs(int m, int half, int p, int npairs) {
    return (int64_t)(m >> 3) * syn_block_elems(npairs) + (int64_t)half * pairs8(npairs) * MB +
           (int64_t)(p >> 3) * (8 * MB) + (m & 7) * 8 + (p & 7);
}

inline int64_t nalm(int lmax, int mmax) {
    return (int64_t)(mmax + 1) * (lmax + 1) - (int64_t)mmax * (mmax + 1) / 2;
}
// healpy m-major index of (l, m)
__host__ __device__ inline int64_t alm_index(int l, int m, int lmax) {
    return (int64_t)m * (2 * lmax + 1 - m) / 2 + l;
}
inline int ilog2(int64_t v) { int r = 0; while ((int64_t(1) << r) < v) ++r; return r; }

// One FFT size class: rings whose inner power-of-two FFT has size M, split
// into rings with h = M (one FFT) and Bluestein rings (h not a power of two).
struct FftClass {
    int M = 0, log2M = 0, threads = 0;
    int nrings_pow2 = 0, nrings_blue = 0;
    int32_t *d_rings_pow2 = nullptr, *d_rings_blue = nullptr;   // ring indices
    int64_t tw_off64 = 0, tw_off32 = 0;   // offsets of this class's FP64 / FP32 twiddle tables
};

// Fused ring launches (ring_fft.hip): one launch runs every FFT size class
// of one workgroup size, for one plan or for every plan ("group") of a
// mixed-geometry batch.  A segment is one plan's rings of one class (or, for
// a batch's FP64 band DFT, all of a group's rings) and owns the blocks from
// block0 (times the call's unit) on.
struct RingSeg {
    const int32_t *rings;          // ring indices of the class (null: band DFT)
    const int64_t *ring_start;
    const int32_t *nphi, *mcut;
    const double *phi0, *scale;    // scale: batches only (one plan: the call's)
    const double2 *rot;
    const void *tw, *V;            // pass twiddles and chirp spectra of the transform's precision
    const int64_t *vofs;
    int64_t pstride;               // phase elements per map
    uint64_t phase_off;            // batches: bytes into the workspace
    int32_t block0, ncls, nmaps, g, M, log2M, mmax, nr;
};
// size_ana / size_syn: blocks (batches) or units of 8 / 64 blocks per map
// (one plan, whose grid scales with the call's map count)
struct RingLaunch {
    int threads = 0, blue = 0, nseg = 0;
    size_t lds = 0;
    int64_t size_ana = 0, size_syn = 0;
    RingSeg *d_ana = nullptr, *d_syn = nullptr;
};
// Per-call arguments: a batch passes maps_tab / mstride_tab (device tables,
// one entry per group) and its workspace, unit 1; one plan passes its maps,
// phase buffer, scale, map count and unit = 8 or 64 blocks per map.
struct RingCall {
    const void *const *maps_tab;
    const int64_t *mstride_tab;
    char *ws;
    void *maps;
    int64_t mstride;
    void *phase;
    const double *scale;
    int nmaps, unit;
};

// Device allocations of a plan.  Plans derived from it with hbx_plan_derive
// share them (read-only tables); the last plan destroyed frees them on their
// device.  They also own the side stream on which a call runs its
// independent FP64 band work (FP32 plans) next to the main work on the
// caller's stream, and the two events that fork and join it (used under
// side_mu).
struct DeviceAllocs {
    int device = 0;
    std::vector<void *> ptrs;
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::mutex side_mu;
    DeviceAllocs() = default;
    DeviceAllocs(const DeviceAllocs &) = delete;
    DeviceAllocs &operator=(const DeviceAllocs &) = delete;
    ~DeviceAllocs();
};
struct AllocRef {
    std::shared_ptr<DeviceAllocs> s = std::make_shared<DeviceAllocs>();
    void push_back(void *ptr) { s->ptrs.push_back(ptr); }
};

struct Plan {
    int device = 0;
    int nrings = 0, npairs = 0, lmax = 0, mmax = 0, nmb = 0, spin = 0, precision = 0;
    int64_t npix = 0, nlm = 0, nrow = 0;   // nrow: padded recursion-row storage
    uint64_t device_bytes = 0;
    // geometry (device)
    int64_t *d_ring_start = nullptr;
    int32_t *d_nphi = nullptr;
    int32_t *d_mcut = nullptr;
    double *d_phi0 = nullptr;              // nullptr when every phi0 is 0
    double *d_scale_ana = nullptr;         // weight_r / nphi_r
    double *d_scale_adj = nullptr;         // weight_r / nphi_r (map2alm adjoint: omega_r)
    int64_t *d_vofs = nullptr;             // per ring offset into chirp spectra (-1: pow2)
    int64_t *d_cofs = nullptr;             // [mmax + 2] prepared synthesis chunks of row m start here
    int64_t syn_chunks = 0;
    double2 *d_rot = nullptr;              // [4][nrings] FP64 rotation steps of the ring kernels
    double2 *d_rot64 = nullptr;            // the same for FP64 analysis transforms of an FP32 plan
    // FP32 plan whose ring analysis runs in FP64 (output rounded once to FP32):
    // spin plans, where B's weak high-l signal sits below the FP32 FFT's
    // absolute error
    bool ring_f64 = false;
    // ring pairs (device)
    int32_t *d_pair_n = nullptr, *d_pair_s = nullptr;
    double *d_px = nullptr;                // cos theta of the north ring of each pair
    // recursion tables (device), rows padded to CHUNK, row m at d_rofs[m]
    int64_t *d_rofs = nullptr;
    double *d_alpha = nullptr, *d_sl = nullptr;
    int32_t *d_lseed = nullptr;            // [mmax+1][npairs]
    double2 *d_seed = nullptr;             // [mmax+1][npairs] (mu_{ls-2}, mu_{ls-1})
    // spin plans: coupling coefficients of the lambda_+- recursion and
    // seeds (mu+_{ls-2}, mu+_{ls-1}, mu-_{ls-2}, mu-_{ls-1})
    double *d_beta = nullptr;
    double4 *d_seed4 = nullptr;
    // FFT tables (device)
    double2 *d_ftw64 = nullptr; float2 *d_ftw32 = nullptr;   // per-class FFT pass twiddles
    double2 *d_V64 = nullptr;  float2 *d_V32 = nullptr;
    int64_t vtotal = 0;
    std::vector<FftClass> classes;
    // ring stage launches: the plan's precision, and (ring_f64) the FP64
    // analysis transforms of an FP32 plan
    std::vector<RingLaunch> fused, fused64;
    AllocRef allocations;
    // tiles (fixed when the plan is created or derived); 0 = size-based default
    int tile_valu_slots = 0;      // VALU analysis ring-pair slots per lane
    int tile_mfma_waves = 0;      // waves per matrix-core analysis workgroup (4, 8, 12)
    int tile_mfma_min_maps = 0;   // FP32 batches of at least this many maps use the matrix cores
    int tile_spin_slots = 0;      // spin analysis ring-pair slots per lane (1, 2, 4)
    int tile_mfma_kind = 0;       // HBX_MFMA_F32 or HBX_MFMA_F16X2 (default)
    int tile_syn_groups = 0;      // 16-map groups per split-fp16 synthesis workgroup (1, 2; 0 = per batch)
};

// launchers implemented in ring_fft.hip / legendre.hip
int plan_build_fft_tables(Plan *p, const std::vector<int32_t> &nphi, const double *phi0);
int launch_ring_analysis(const Plan *p, int nmaps, const void *maps, int map_dtype,
                         int64_t map_stride, void *phase, const double *scale,
                         hipStream_t s);
int launch_ring_synthesis(const Plan *p, int nmaps, const void *phase, void *maps,
                          int map_dtype, int64_t map_stride, const double *scale,
                          hipStream_t s);

// a batch's ring launches over every group's classes, largest workgroups
// first, and (FP32 batches) its band DFT launch; device tables are appended
// to allocs
int build_ring_launches(const std::vector<Plan> &plans, const std::vector<int> &nmaps,
                        const std::vector<uint64_t> &phase_off, std::vector<RingLaunch> &out,
                        RingLaunch &band, std::vector<void *> &allocs);
int launch_multi_ring_analysis(const RingLaunch &L, int precision, int map_dtype, const void *const *maps_tab,
                               const int64_t *mstride_tab, char *ws, hipStream_t s);
int launch_multi_ring_synthesis(const RingLaunch &L, int precision, int map_dtype, void *const *maps_tab,
                                const int64_t *mstride_tab, char *ws, hipStream_t s);
int launch_multi_ring_band(const RingLaunch &L, int map_dtype, const void *const *maps_tab,
                           const int64_t *mstride_tab, char *ws, hipStream_t s);
// recursion rows: row m holds l = m .. m + len - 1 (len a multiple of CHUNK);
// spin plans start their chunks at l0 = max(m, |s|) and get longer rows
int plan_build_recursion(Plan *p, const std::vector<double> &pair_cos, const std::vector<double> &pair_sin,
                         double tol_log2);
// spin.hip
int plan_build_spin(Plan *p, const std::vector<double> &pair_cos, const std::vector<double> &pair_sin,
                    double tol_log2);
int spin_groups(const Plan *p);
int spin_slots_of(const Plan *p);
int launch_spin_analysis_band(const Plan *p, int npol, const void *band_phase, void *alm, int64_t alm_stride,
                              void *band_alm, void *band_partial, int adjoint, hipStream_t s);
int launch_band_to_f32(const Plan *p, int nrows, const void *band_alm, void *alm, int64_t alm_stride, hipStream_t s);
bool spin_uses_mfma(const Plan *p, int npol);
int spin_mfma_groups(const Plan *p);
int launch_spin_analysis_mfma(const Plan *p, int npol, const void *phase, void *alm, int64_t astride, void *partial,
                              int32_t *lfirst, int adjoint, hipStream_t s);
int launch_spin_analysis(const Plan *p, int npol, const void *phase, void *alm, int64_t alm_stride, void *partial,
                         int adjoint, hipStream_t s);
int launch_spin_prescale(const Plan *p, int npol, const void *alm, int64_t alm_stride, void *scaled, int adjoint,
                         hipStream_t s);
int launch_spin_synthesis(const Plan *p, int npol, const void *scaled, void *phase, hipStream_t s);
int launch_leg_analysis(const Plan *p, int nmaps, const void *phase, void *alm,
                        int64_t alm_stride, void *partial, int adjoint, hipStream_t s);
int analysis_groups(const Plan *p, bool f64);   // ring-pair groups (partial sums) per m
int band_groups(const Plan *p);                 // the same for the FP64 band of FP32 plans
int launch_prescale(const Plan *p, int nmaps, const void *alm, int64_t alm_stride,
                    void *scaled, int adjoint, hipStream_t s);
int launch_leg_synthesis(const Plan *p, int nmaps, const void *scaled, void *phase,
                         hipStream_t s);
// ws: syn_prep_bytes (the split-fp16 kernel's prepared operands and scales)
int launch_leg_synthesis_mfma(const Plan *p, int nmaps, const void *alm, int64_t alm_stride,
                              void *phase, void *ws, int adjoint, hipStream_t s);
int mfma_kind_of(const Plan *p);   // HBX_MFMA_F32 or HBX_MFMA_F16X2
int plan_build_syn_chunks(Plan *p);
uint64_t syn_prep_bytes(const Plan *p, int nmaps);   // split-fp16 synthesis workspace
int launch_leg_analysis_mfma(const Plan *p, int nmaps, const void *phase, void *alm, int64_t alm_stride,
                             void *partial, int32_t *lfirst, int adjoint, hipStream_t s);
// lfirst (may be null): [m][group] first l whose partial sum was written;
// below it the group's partial is zero and is not read
int launch_group_reduce(const Plan *p, int nmaps, void *partial, void *alm, int64_t alm_stride,
                        int ngroups, int64_t nlm_stride, int mbeg, int mend, bool f64, int adjoint,
                        hipStream_t s, const int32_t *lfirst = nullptr);
int launch_leg_analysis_band(const Plan *p, int nmaps, const void *band_phase, void *alm,
                             int64_t alm_stride, void *band_alm, void *band_partial, int adjoint,
                             hipStream_t s);
int launch_ring_band(const Plan *p, int nmaps, const void *maps, int map_dtype, int64_t map_stride,
                     void *band_phase, const double *scale, hipStream_t s);
int mfma_analysis_groups(const Plan *p);
int band_mfma_groups(const Plan *p);
// FP64 band of FP32 batches on the FP64 matrix cores: band phase -> band alm (FP64, scaled)
int launch_band_analysis_mfma(const Plan *p, int nmaps, const void *band, void *band_alm, void *band_partial,
                              int adjoint, hipStream_t s);
int mfma_analysis_waves(const Plan *p);
int prepare_mfma_kernels();
// by default, batches at least this large use the matrix-core Legendre
// kernels (FP32 plans): scalar 3 maps, spin 6 maps (3 Q/U pairs).  Map2alm +
// alm2map at nside 2048, matrix / VALU: 2 maps 19.4 / 19.1 ms, 3 maps 20.2 /
// 28.1 ms; spin 1 pair 53.0 / 27.5 ms, 2 pairs 56.1 / 53.3 ms, 4 pairs 65 /
// 106 ms (a matrix tile serves 16 maps or 4 pairs at nearly the same cost)
constexpr int MFMA_MIN_MAPS = 3;
constexpr int SPIN_MFMA_MIN_MAPS = 6;
inline int mfma_min_maps(const Plan *p) {
    if (p->tile_mfma_min_maps > 0) return p->tile_mfma_min_maps;
    return p->spin != 0 ? SPIN_MFMA_MIN_MAPS : MFMA_MIN_MAPS;
}
int analysis_slots_of(const Plan *p, bool f64);   // effective VALU analysis slots
int launch_alm2cl(int lmax, int mmax, int nmaps, const void *alm, int dt,
                  int64_t alm_stride, double *cl, int64_t cl_stride, hipStream_t s);

int64_t phase_elems_per_map(const Plan *p);   // complex elements
// FP32 plans: FP64 band phase [map][m/8 < NMB_LO][ring][m%8] and band alm (m < M_LO)
inline int64_t band_phase_elems(const Plan *p) { return (int64_t)NMB_LO * p->nrings * MB; }
// coefficients in rows m < min(M_LO, mmax+1) (rows are contiguous in m-major order)
inline int64_t band_nlm(const Plan *p) {
    const int mt = p->mmax + 1 < M_LO ? p->mmax + 1 : M_LO;
    return alm_index(mt, mt, p->lmax);
}
int64_t scaled_elems_per_map(const Plan *p);  // complex elements

// hbx_abi.hip: phase buffer of nmaps maps (FP32 plans: plus the FP64 band
// phase, which starts at band_phase_ptr) and the ring stage of an analysis
// (FP32 plans: the band ring DFT on the plan's side stream)
uint64_t phase_total_bytes(const Plan *p, int nmaps);
void *band_phase_ptr(const Plan *p, int nmaps, const void *phase);
uint64_t band_phase_offset(const Plan *p, int nmaps);   // bytes from the phase buffer's start
int ring_analysis(const Plan *p, int nmaps, const void *maps, int map_dtype, int64_t map_stride, void *phase,
                  const double *scale, hipStream_t s);

}  // namespace hbx

struct hbx_plan {
    hbx::Plan impl;
};
