/*
 * hbx.h - C ABI of libhbx.so, the MI355X (gfx950) spherical-harmonic transform
 * library behind hellblast_amd.
 *
 * Every entry point returns 0 (HBX_OK) or a negative hbx_status; the message
 * of the last failure on the calling thread is returned by hbx_last_error().
 * Data pointers passed to execute functions are caller-owned DEVICE pointers
 * (contiguous, complex arrays interleaved re/im like torch complex64/128).
 * Execute functions never allocate and never synchronise: they enqueue on the
 * given hipStream_t (NULL = default stream) and are safe to capture in a HIP
 * graph.  Analyses with FP32 plans run their FP64 low-m band on a stream of
 * the plan, forked from the given stream and joined back to it: work the
 * caller enqueues on the given stream after the call runs after the band
 * (everything stays on the given stream while it is being captured; FP32
 * mixed-geometry batches fork their band also then, see hbx_multi_create).
 * Plans are immutable after creation (tiles included: another tile
 * choice is another plan, made by hbx_plan_derive) and bound to the device
 * that was current when they were created; one plan may be used from several
 * threads and streams at once.
 *
 * Reference interfaces replaced (file:line in Luce-Morningstar/hellblast):
 *   hbx_plan_create      - per-call setup in sht_raw.py:30-45 (ring FFT prep,
 *                          x = cos(theta)) and _compute_legendre_block
 *                          sht_raw.py:16-24 / legendre.cu:2-35 (moved to plan
 *                          time as recursion tables and switch-on seeds)
 *   hbx_map2alm          - sht_raw.map2alm_gpu        sht_raw.py:30-57
 *                          sht_fallback.map2alm_gpu   sht_fallback.py:36-65
 *   hbx_alm2map          - isht.alm2map_gpu           isht.py:5-26
 *   hbx_alm2cl           - alm_ops.alm2cl             alm_ops.py:19-32
 *   hbx_ang2pix_ring     - coords_ring ang2pix_ring_with_rings
 *                          coords_ring.cu:9-32 (+ ring search, coords_ring.py:18-24)
 *   hbx_pix2ang_ring     - coords_ring pix2ang_ring_batch coords_ring.cu:33-54
 *   hbx_map2alm_spin     - spin.map2alm_spin_gpu      spin.py:32-71
 *                          spin1.map2alm_spin1_gpu    spin1.py:5-38
 *   hbx_alm2map_spin     - spin.qu_from_eb_gpu        spin.py:80-115
 *                          spin1.alm2map_spin1_gpu    spin1.py:40-69
 *   hbx_alm2map_adjoint, hbx_map2alm_adjoint, hbx_alm2map_spin_adjoint,
 *   hbx_map2alm_spin_adjoint - no reference counterpart (the reference has no
 *                          gradients); they are the exact transposes used by
 *                          hellblast_amd.autograd.
 */
#ifndef HBX_H
#define HBX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HBX_ABI_VERSION 3

typedef enum hbx_status {
    HBX_OK = 0,
    HBX_EINVAL = -1,       /* bad argument (shape, range, null pointer)      */
    HBX_ENOMEM = -2,       /* device allocation failed at plan creation      */
    HBX_EDEVICE = -3,      /* HIP runtime error / no device                  */
    HBX_EWORKSPACE = -4,   /* workspace smaller than hbx_workspace_bytes()   */
    HBX_EUNSUPPORTED = -5, /* geometry or size outside what the kernels do   */
    HBX_EPLAN = -6         /* plan used on another device or wrong kind      */
} hbx_status;

typedef enum hbx_precision {
    HBX_F64 = 0,           /* FP64 data, FP64 recursion, FP64 contraction    */
    HBX_F32 = 1            /* FP32 data + contraction, FP64 recursion/seeds  */
} hbx_precision;

typedef enum hbx_dtype {
    HBX_DT_F64 = 0,
    HBX_DT_F32 = 1
} hbx_dtype;

/* Ring geometry, all arrays in HOST memory; the plan copies what it needs.
 * Rings r and nrings-1-r must be mirror images (equal nphi, cos theta of
 * opposite sign); an odd middle ring must sit on the equator.  Every nphi
 * must be even.  This holds for HEALPix RING ordering.                      */
typedef struct hbx_geom {
    int32_t nrings;
    const int64_t *ring_start; /* nrings+1 offsets into the flat map          */
    const int32_t *nphi;       /* pixels per ring                             */
    const double *cos_theta;   /* cos theta per ring                          */
    const double *sin_theta;   /* sin theta per ring (>= 0)                   */
    const double *phi0;        /* longitude of pixel 0 per ring, NULL = 0     */
    const double *weight;      /* analysis weight per ring (multiplies
                                  DFT/nphi, e.g. 2*pi*dz), NULL = 1          */
    const int32_t *mcut;       /* ring r contributes to m < mcut[r] with bin
                                  m mod nphi; NULL = nphi (reference rule)   */
} hbx_geom;

typedef struct hbx_plan hbx_plan;

typedef struct hbx_plan_info {
    int32_t nrings, npairs, lmax, mmax, precision, spin;
    int64_t npix, nalm;
    uint64_t device_bytes;     /* device memory held by the plan             */
    int32_t fp64_band_m;       /* FP32 scalar plans: rows m < fp64_band_m are
                                  analysed in FP64 (0 otherwise)             */
    int32_t reserved;
} hbx_plan_info;

int hbx_abi_version(void);
const char *hbx_last_error(void);

/* Host-only validation of a geometry (no device needed). */
int hbx_validate_geom(const hbx_geom *g, int32_t lmax, int32_t mmax);

/* spin: 0 (scalar) or 1 .. HBX_MAX_SPIN (spin-weighted, spin <= lmax; the
 * reference's spin = 2 and spin = 1 transforms).  A spin plan serves only the
 * hbx_*_spin entry points and a spin-0 plan only the others (HBX_EPLAN).
 * flags: reserved, pass 0.                                                  */
#define HBX_MAX_SPIN 8
int hbx_plan_create(const hbx_geom *g, int32_t lmax, int32_t mmax, int32_t spin,
                    int32_t precision, uint32_t flags, hbx_plan **out);
int hbx_plan_destroy(hbx_plan *plan);
int hbx_plan_get_info(const hbx_plan *plan, hbx_plan_info *info);
int hbx_workspace_bytes(const hbx_plan *plan, int32_t nmaps, uint64_t *bytes);

/* Tiles (no reference counterpart; the autotuner's hook).  0 = the
 * size-based default.  hbx_plan_derive makes a new plan with the given tiles
 * that shares the device tables of base (no device allocation; base and the
 * derived plan may be destroyed in any order, the tables are freed with the
 * last of them).  Tiles change the workspace size and the summation order
 * (results agree to rounding).  get_tiles returns the effective values.
 *   valu_slots     VALU Legendre analysis: ring-pair slots of 64 pairs per lane
 *                  (power of two <= 16; FP64 kernels clamp to 8)
 *   mfma_waves     matrix-core analysis: waves (64 ring pairs each) per
 *                  workgroup, 4 or 8; or 12: 512 ring pairs per workgroup,
 *                  8 MFMA waves fed by 4 lambda-recursion waves one chunk
 *                  ahead (double-buffered LDS tile)
 *   mfma_min_maps  FP32 batches of at least this many maps use the
 *                  matrix-core Legendre kernels (default 3; spin plans count
 *                  the Q and U maps, default 6 = 3 Q/U pairs)
 *   spin_slots     VALU spin Legendre analysis (FP64 plans, and FP32 batches
 *                  below mfma_min_maps): ring-pair slots per lane, 1, 2, 4
 *   mfma_kind      matrix-core operand format of FP32 batches:
 *                  HBX_MFMA_F32 (FP32 instructions) or HBX_MFMA_F16X2
 *                  (each FP32 operand as an fp16 hi/lo pair, three fp16
 *                  products per FP32 product, FP32 accumulation; the
 *                  default); with HBX_MFMA_F16X2, mfma_waves is 4, 8 or 16
 *   mfma_groups    groups of 16 maps per split-fp16 workgroup: 1, or 2
 *                  (each lambda value feeds twice the matrix work, at two
 *                  thirds of the occupancy in the synthesis; in the analysis
 *                  16-wave workgroups with mfma_waves 8, 8-wave workgroups
 *                  with mfma_waves 4); 0 (the default): the synthesis takes
 *                  2 for batches of an even number of 16-map groups, else
 *                  1, and the analysis 2 for such batches where one 256-pair
 *                  workgroup covers every ring pair (nside <= 128), else 1.
 *                  The results do not depend on it (bit for bit). */
#define HBX_MFMA_F32 1
#define HBX_MFMA_F16X2 2
typedef struct hbx_tiles {
    int32_t valu_slots, mfma_waves, mfma_min_maps, spin_slots, mfma_kind, mfma_groups;
} hbx_tiles;
int hbx_plan_derive(const hbx_plan *base, const hbx_tiles *tiles, hbx_plan **out);
int hbx_plan_get_tiles(const hbx_plan *plan, hbx_tiles *tiles);

/* Scalar transforms.  maps: nmaps maps of plan npix values (dtype
 * map_dtype), map_stride elements apart.  alm: nmaps m-major alm arrays
 * (idx = m(2 lmax+1-m)/2 + l) of complex128 (HBX_F64 plans) or complex64
 * (HBX_F32 plans), alm_stride complex elements apart.                       */
int hbx_map2alm(const hbx_plan *plan, int32_t nmaps, const void *maps,
                int32_t map_dtype, int64_t map_stride, void *alm,
                int64_t alm_stride, void *ws, uint64_t ws_bytes, void *stream);
int hbx_alm2map(const hbx_plan *plan, int32_t nmaps, const void *alm,
                int64_t alm_stride, void *maps, int32_t map_dtype,
                int64_t map_stride, void *ws, uint64_t ws_bytes, void *stream);
/* hbx_map2alm with no workspace argument that allocates nothing: its
 * workspace is a static device array of HBX_STATIC_SCRATCH_BYTES, which each
 * GPU holds from the first call on it until the process ends.  No
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  A plan and batch whose
 * hbx_workspace_bytes() exceeds the scratch get HBX_EUNSUPPORTED (64 MiB
 * holds e.g. one FP64 map at nside 512, lmax 1023).  All calls on one GPU
 * share the scratch: calls that overlap on one GPU (other streams or
 * threads) race.                                                           */
#define HBX_STATIC_SCRATCH_BYTES (64u << 20)
int hbx_map2alm_static(const hbx_plan *plan, int32_t nmaps, const void *maps,
                       int32_t map_dtype, int64_t map_stride, void *alm,
                       int64_t alm_stride, void *stream);
/* Transposes: adjoint of alm2map (c_m-weighted unweighted analysis) and of
 * map2alm (weighted synthesis of g/c_m), PyTorch complex-gradient
 * convention grad = dL/dRe + i dL/dIm.                                      */
int hbx_alm2map_adjoint(const hbx_plan *plan, int32_t nmaps, const void *maps,
                        int32_t map_dtype, int64_t map_stride, void *alm,
                        int64_t alm_stride, void *ws, uint64_t ws_bytes,
                        void *stream);
int hbx_map2alm_adjoint(const hbx_plan *plan, int32_t nmaps, const void *alm,
                        int64_t alm_stride, void *maps, int32_t map_dtype,
                        int64_t map_stride, void *ws, uint64_t ws_bytes,
                        void *stream);

/* Spin-weighted transforms, healpy E/B convention:
 *   Q + iU = sum_lm  s_a_lm  sY_lm,  s_a_lm = -(E_lm + i B_lm),
 *   Q - iU = sum_lm -s_a_lm -sY_lm, -s_a_lm = -(-1)^s (E_lm - i B_lm),
 * sY_lm as in Goldberg et al. (1967) (2Y_20 = 1/4 sqrt(15/2pi) sin^2 theta).
 * npol polarisation pairs: maps 2b and 2b+1 are Q and U of pair b (map_stride
 * apart), alm rows 2b and 2b+1 are E and B (alm_stride apart, m-major).
 * Entries l < max(m, spin) are written as 0.  Workspace: hbx_workspace_bytes
 * with nmaps = npol.  npol <= 32767.
 *   hbx_map2alm_spin  (Q, U) -> (E, B)  eb_from_qu_gpu spin.py:73-78 (with
 *                     map2alm_spin_gpu spin.py:32-71); spin1.py:5-38
 *   hbx_alm2map_spin  (E, B) -> (Q, U)  qu_from_eb_gpu spin.py:80-115;
 *                     spin1.py:40-69
 *   hbx_alm2map_spin_adjoint, hbx_map2alm_spin_adjoint: transposes, as for
 *                     the scalar pair.                                      */
int hbx_map2alm_spin(const hbx_plan *plan, int32_t npol, const void *maps,
                     int32_t map_dtype, int64_t map_stride, void *alm,
                     int64_t alm_stride, void *ws, uint64_t ws_bytes,
                     void *stream);
int hbx_alm2map_spin(const hbx_plan *plan, int32_t npol, const void *alm,
                     int64_t alm_stride, void *maps, int32_t map_dtype,
                     int64_t map_stride, void *ws, uint64_t ws_bytes,
                     void *stream);
int hbx_alm2map_spin_adjoint(const hbx_plan *plan, int32_t npol,
                             const void *maps, int32_t map_dtype,
                             int64_t map_stride, void *alm, int64_t alm_stride,
                             void *ws, uint64_t ws_bytes, void *stream);
int hbx_map2alm_spin_adjoint(const hbx_plan *plan, int32_t npol,
                             const void *alm, int64_t alm_stride, void *maps,
                             int32_t map_dtype, int64_t map_stride, void *ws,
                             uint64_t ws_bytes, void *stream);

/* Stage entry points (the two halves of every transform).  The phase
 * buffer holds F_r(m) / G_r(m) for every ring and m <= mmax as complex
 * (precision of the plan), maps S = nmb * 2 * pairs8 * 8 elements apart
 * (nmb = ceil((mmax+1)/8), npairs = ceil(nrings/2), pairs8 = npairs rounded
 * up to a multiple of 8), in m-blocks of 8 m laid out by direction so that
 * each stage's writer fills whole 64-byte lines:
 *   analysis (ring -> Legendre): [m/8][ring][m%8], blocks nrings * 8 apart;
 *   synthesis (Legendre -> ring): [m/8][half][pair/8][m%8][pair%8], blocks
 *     2 * pairs8 * 8 apart, half 0 (pairs8 * 8 elements) holding ring p and
 *     half 1 ring nrings-1-p of ring pair p (ring r >= npairs is half 1 of
 *     pair nrings-1-r).
 * hbx_phase_bytes() gives the size.  The stages let a caller time or
 * overlap them, or exchange the phase buffer between devices
 * (ring-parallel FFT stage, m-parallel Legendre stage).
 *   hbx_ring_analysis      - per-ring FFT loop      sht_raw.py:38-43
 *   hbx_legendre_analysis  - per-m Legendre loop    sht_raw.py:46-56
 *   hbx_legendre_synthesis - per-m l-sum            isht.py:11-23
 *   hbx_ring_synthesis     - per-ring inverse FFT   isht.py:16-25
 * The Legendre stages need ws of hbx_workspace_bytes() bytes.             */
int hbx_phase_bytes(const hbx_plan *plan, int32_t nmaps, uint64_t *bytes);
int hbx_ring_analysis(const hbx_plan *plan, int32_t nmaps, const void *maps,
                      int32_t map_dtype, int64_t map_stride, void *phase,
                      uint64_t phase_bytes, void *stream);
int hbx_legendre_analysis(const hbx_plan *plan, int32_t nmaps,
                          const void *phase, uint64_t phase_bytes, void *alm,
                          int64_t alm_stride, void *ws, uint64_t ws_bytes,
                          void *stream);
int hbx_legendre_synthesis(const hbx_plan *plan, int32_t nmaps,
                           const void *alm, int64_t alm_stride, void *phase,
                           uint64_t phase_bytes, void *ws, uint64_t ws_bytes,
                           void *stream);
int hbx_ring_synthesis(const hbx_plan *plan, int32_t nmaps, const void *phase,
                       uint64_t phase_bytes, void *maps, int32_t map_dtype,
                       int64_t map_stride, void *stream);

/* C_l = (|a_l0|^2 + 2 sum_{m>=1} |a_lm|^2)/(2l+1) for nmaps m-major alm of
 * complex128 (alm_dtype HBX_DT_F64) or complex64 (HBX_DT_F32); cl is FP64
 * [nmaps][lmax+1] with cl_stride elements between maps.                     */
int hbx_alm2cl(int32_t lmax, int32_t mmax, int32_t nmaps, const void *alm,
               int32_t alm_dtype, int64_t alm_stride, double *cl,
               int64_t cl_stride, void *stream);

/* RING pixel <-> angle (device arrays; tables as in HealpixRingTables:
 * z_edges[nrings+1] decreasing from 1 to -1, ring_start[nrings+1], nphi,
 * phi0 = phi of pixel 0 of each ring).  ang2pix: ring = the z-edge cell of
 * cos(theta), pixel = nearest centre phi0_r + j 2 pi / nphi_r.  pix2ang:
 * (theta_centers[r], phi0_r + j 2 pi / nphi_r wrapped to [0, 2 pi)); an
 * index outside [0, ring_start[nrings]) gives NaN angles.                   */
int hbx_ang2pix_ring(int64_t n, const double *theta, const double *phi,
                     int32_t nrings, const double *z_edges,
                     const int64_t *ring_start, const int32_t *nphi,
                     const double *phi0, int64_t *ipix, void *stream);
int hbx_pix2ang_ring(int64_t n, const int64_t *ipix, int32_t nrings,
                     const double *theta_centers, const int64_t *ring_start,
                     const int32_t *nphi, const double *phi0, double *theta,
                     double *phi, void *stream);

/* Mixed-geometry batches in persistent launches (no reference counterpart:
 * the reference transforms one map per call, sht_raw.py:30-57 and
 * isht.py:5-26; this is the batched form of those two entry points).  A
 * batch holds ngroups spin-0 plans of one precision (each its own geometry,
 * lmax and mmax) with nmaps[g] maps each; it copies what it needs, so the
 * plans may be destroyed afterwards.  hbx_multi_map2alm / hbx_multi_alm2map
 * transform every map of every group in a fixed number of launches: the
 * ring stage runs in fused launches over the rings of every group (one per
 * FFT workgroup size), and each Legendre direction is ONE persistent launch
 * whose workgroups take (group, m-block, ring-pair block, map block) tiles
 * from work queues, so the work of large and small maps spreads over the
 * whole GPU in one kernel.  FP32 batches compute m < 8 in FP64 (band DFT
 * and one persistent FP64 launch) on a stream of the batch, forked from the
 * given stream after the call's first kernel and joined back to it before
 * the call returns, by two events under the batch's lock; this also holds
 * while the given stream is being captured (the fork and join become graph
 * edges).  The ring stages and Legendre tiles run the per-plan VALU
 * kernels' arithmetic: results
 * equal hbx_map2alm / hbx_alm2map of each group with those kernels
 * (valu_slots equal to the plan's, mfma_min_maps above the batch) bit for
 * bit.  Per group g: maps[g] (nmaps[g] maps, map_stride[g] apart, dtype
 * map_dtype) and alm[g] (alm_stride[g] apart).  ws: at least
 * hbx_multi_workspace_bytes(); it also holds the call's work-queue counters,
 * so calls that overlap in time need their own workspaces.  Calls never
 * allocate or synchronise and can be captured in a HIP graph.
 * Tiles (0 = default):
 *   valu_slots      analysis ring-pair slots of 64 pairs per lane: 1, 2, 4, 8
 *                   (default 4 for FP32, 8 for FP64)
 *   syn_maps        maps of one group per synthesis tile sharing each
 *                   recursion step: 1, 2, 4 (default 4)
 *   tiles_per_wave  analysis tiles per resident wave the m ranges are cut
 *                   for (default 8; larger = smaller tiles, finer balance) */
#define HBX_MULTI_MAX_GROUPS 16
typedef struct hbx_multi hbx_multi;
typedef struct hbx_multi_tiles {
    int32_t valu_slots, syn_maps, tiles_per_wave, reserved;
} hbx_multi_tiles;
typedef struct hbx_multi_info {
    int32_t ngroups, precision, valu_slots, syn_maps, tiles_per_wave;
    int32_t analysis_tiles, synthesis_tiles, analysis_workgroups, synthesis_workgroups, reserved;
    uint64_t workspace_bytes;
} hbx_multi_info;
int hbx_multi_create(const hbx_plan *const *plans, const int32_t *nmaps, int32_t ngroups,
                     const hbx_multi_tiles *tiles, hbx_multi **out);
int hbx_multi_destroy(hbx_multi *batch);
int hbx_multi_workspace_bytes(const hbx_multi *batch, uint64_t *bytes);
int hbx_multi_get_info(const hbx_multi *batch, hbx_multi_info *info);
int hbx_multi_map2alm(const hbx_multi *batch, const void *const *maps, const int64_t *map_stride,
                      int32_t map_dtype, void *const *alm, const int64_t *alm_stride, void *ws,
                      uint64_t ws_bytes, void *stream);
int hbx_multi_alm2map(const hbx_multi *batch, const void *const *alm, const int64_t *alm_stride,
                      void *const *maps, const int64_t *map_stride, int32_t map_dtype, void *ws,
                      uint64_t ws_bytes, void *stream);

/* Kernel timers (diagnostics; no reference counterpart).  hbx_timer_arm(mask)
 * arms the kernels whose bits (1 << HBX_TIMER_*) are set and clears their
 * records (mask 0 disarms all).  While armed, each launch of such a kernel is
 * bracketed by HIP events recorded on the stream it is launched on (up to
 * 256 launches per kernel; later launches are not recorded).
 * hbx_timer_read() waits for one kernel's recorded events and returns the
 * number of recorded launches and their summed duration in ms.  Do not arm
 * timers while a stream is being captured into a HIP graph.
 *   HBX_TIMER_LEGENDRE_ANALYSIS  - main Legendre analysis kernel (matrix-core
 *                                  kernel for FP32 batches; excludes the FP64
 *                                  low-m band and partial-sum reductions)
 *   HBX_TIMER_LEGENDRE_SYNTHESIS - main Legendre synthesis kernel
 *   HBX_TIMER_RING_ANALYSIS      - all ring-FFT launches of one ring analysis
 *   HBX_TIMER_RING_SYNTHESIS     - all ring-FFT launches of one ring synthesis
 *   HBX_TIMER_SPIN_ANALYSIS, HBX_TIMER_SPIN_SYNTHESIS - spin Legendre kernels */
#define HBX_TIMER_LEGENDRE_ANALYSIS 1
#define HBX_TIMER_LEGENDRE_SYNTHESIS 2
#define HBX_TIMER_RING_ANALYSIS 3
#define HBX_TIMER_RING_SYNTHESIS 4
#define HBX_TIMER_SPIN_ANALYSIS 5
#define HBX_TIMER_SPIN_SYNTHESIS 6
int hbx_timer_arm(uint32_t mask);
int hbx_timer_read(int32_t kernel, int32_t *launches, double *total_ms);

#ifdef __cplusplus
}
#endif

#endif /* HBX_H */
