// RING-ordered pixel <-> angle over the ring tables.
//
// Reference: ang2pix_ring_with_rings / pix2ang_ring_batch (coords_ring.cu:9-54)
// driven by RingMapperGPU (coords_ring.py:18-56).  The reference's rule is
// kept: the ring of (theta, phi) is the z-edge cell holding z = cos(theta)
// (edges are the midpoints between ring centres, coords_ring.py:18-24), the
// pixel is the nearest centre in phi, phi_c = phi0_r + j 2 pi / nphi_r.  The
// ring search is a binary search here (the reference scans ring_start
// linearly from ring 0 for every pixel); a pixel index outside [0, npix)
// gives NaN angles instead of reading past the tables.
#include "hbx_device.h"
#include "hbx_internal.h"

// separate multiplies and adds (no FMA contraction): the angles are the
// correctly rounded results of each step, as in the NumPy restatement
#pragma clang fp contract(off)

namespace hbx {

__device__ __forceinline__ double two_pi() { return 6.28318530717958647692528676655900577; }

// z_edges[0..nr] decreasing from 1 to -1: r = #{k : z_edges[k] >= z} - 1, in [0, nr)
__device__ __forceinline__ int ring_of_z(double z, const double *__restrict__ z_edges, int nr) {
    int lo = 0, hi = nr + 1;                 // count of edges >= z lies in [lo, hi]
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (z_edges[mid] >= z) lo = mid + 1;
        else hi = mid;
    }
    return min(max(lo - 1, 0), nr - 1);
}

__global__ void ang2pix_ring_kernel(int64_t n, const double *__restrict__ theta, const double *__restrict__ phi,
                                    int nr, const double *__restrict__ z_edges,
                                    const int64_t *__restrict__ ring_start, const int32_t *__restrict__ nphi,
                                    const double *__restrict__ phi0, int64_t *__restrict__ ipix) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int r = ring_of_z(cos(theta[i]), z_edges, nr);
        const int npr = nphi[r];
        const double dphi = two_pi() / (double)npr;
        double d = fmod(phi[i] - phi0[r], two_pi());
        if (d < 0.0) d += two_pi();
        long long j = llround(d / dphi);
        if (j >= npr) j -= npr;
        ipix[i] = ring_start[r] + j;
    }
}

__global__ void pix2ang_ring_kernel(int64_t n, const int64_t *__restrict__ ipix, int nr,
                                    const double *__restrict__ theta_centers,
                                    const int64_t *__restrict__ ring_start, const int32_t *__restrict__ nphi,
                                    const double *__restrict__ phi0, double *__restrict__ theta,
                                    double *__restrict__ phi) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = ipix[i];
        if (p < 0 || p >= ring_start[nr]) {
            theta[i] = __builtin_nan("");
            phi[i] = __builtin_nan("");
            continue;
        }
        int lo = 0, hi = nr - 1;             // last r with ring_start[r] <= p
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ring_start[mid] <= p) lo = mid;
            else hi = mid - 1;
        }
        const int r = lo;
        const double dphi = two_pi() / (double)nphi[r];
        double f = phi0[r] + (double)(p - ring_start[r]) * dphi;
        if (f >= two_pi()) f -= two_pi() * floor(f / two_pi());
        theta[i] = theta_centers[r];
        phi[i] = f;
    }
}

static unsigned coord_blocks(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 65536); }

}  // namespace hbx

using namespace hbx;

extern "C" {

int hbx_ang2pix_ring(int64_t n, const double *theta, const double *phi, int32_t nrings, const double *z_edges,
                     const int64_t *ring_start, const int32_t *nphi, const double *phi0, int64_t *ipix,
                     void *stream) {
    if (n < 0 || nrings < 1) return set_error(HBX_EINVAL, "need n >= 0 and nrings >= 1");
    if (n == 0) return HBX_OK;
    if (!theta || !phi || !z_edges || !ring_start || !nphi || !phi0 || !ipix)
        return set_error(HBX_EINVAL, "data pointer is NULL");
    hipLaunchKernelGGL(ang2pix_ring_kernel, dim3(coord_blocks(n)), dim3(256), 0, (hipStream_t)stream, n, theta, phi,
                       (int)nrings, z_edges, ring_start, nphi, phi0, ipix);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "ang2pix_ring_kernel launch");
    return HBX_OK;
}

int hbx_pix2ang_ring(int64_t n, const int64_t *ipix, int32_t nrings, const double *theta_centers,
                     const int64_t *ring_start, const int32_t *nphi, const double *phi0, double *theta, double *phi,
                     void *stream) {
    if (n < 0 || nrings < 1) return set_error(HBX_EINVAL, "need n >= 0 and nrings >= 1");
    if (n == 0) return HBX_OK;
    if (!ipix || !theta_centers || !ring_start || !nphi || !phi0 || !theta || !phi)
        return set_error(HBX_EINVAL, "data pointer is NULL");
    hipLaunchKernelGGL(pix2ang_ring_kernel, dim3(coord_blocks(n)), dim3(256), 0, (hipStream_t)stream, n, ipix,
                       (int)nrings, theta_centers, ring_start, nphi, phi0, theta, phi);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "pix2ang_ring_kernel launch");
    return HBX_OK;
}

}  // extern "C"
