// hbx_map2alm_static: hbx_map2alm with its workspace in a file-scope device
// array instead of a caller-provided one, for callers that cannot allocate
// device memory.  Each GPU the library runs on holds one copy of the array
// (HBX_STATIC_SCRATCH_BYTES) from the first call on it until the process
// ends; every call on that GPU shares it, so calls that overlap on one GPU
// (other streams, other threads) race.  The kernels write every workspace
// entry they read, so nothing is cleared between calls.
#include <mutex>
#include <string>

#include "hbx_internal.h"

namespace hbx {

static __device__ __align__(256) unsigned char scratch[HBX_STATIC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[MAX_DEVICES];   // device address of scratch per GPU (nullptr: not looked up)

// The address of scratch on the current device, looked up once per device.
// (The array's name in host code is only a handle for the lookup, not a
// device address.)
static int scratch_on(int dev, void **out) {
    if (dev < 0 || dev >= MAX_DEVICES) return set_error(HBX_EDEVICE, "device index out of range for scratch");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_error(e, "hipGetSymbolAddress(scratch)");
        if (!ptr) return set_error(HBX_EDEVICE, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return HBX_OK;
}

}  // namespace hbx

using namespace hbx;

extern "C" int hbx_map2alm_static(const hbx_plan *plan, int32_t nmaps, const void *maps, int32_t map_dtype,
                                  int64_t map_stride, void *alm, int64_t alm_stride, void *stream) {
    if (!plan) return set_error(HBX_EINVAL, "plan is NULL");
    if (nmaps < 0) return set_error(HBX_EINVAL, "nmaps must be >= 0");
    uint64_t need = 0;
    int rc = hbx_workspace_bytes(plan, nmaps, &need);
    if (rc) return rc;
    if (need > (uint64_t)HBX_STATIC_SCRATCH_BYTES)
        return set_error(HBX_EUNSUPPORTED, "hbx_map2alm_static: this plan and batch need " + std::to_string(need) +
                                               " workspace bytes, above the static scratch of " +
                                               std::to_string((uint64_t)HBX_STATIC_SCRATCH_BYTES) +
                                               "; use hbx_map2alm with a workspace");
    if (nmaps == 0) return HBX_OK;
    int dev = -1;
    HBX_HIP(hipGetDevice(&dev));
    if (dev != plan->impl.device)
        return set_error(HBX_EPLAN, "plan was created on device " + std::to_string(plan->impl.device) +
                                        " but device " + std::to_string(dev) + " is current");
    void *ws = nullptr;
    if ((rc = scratch_on(dev, &ws))) return rc;
    return hbx_map2alm(plan, nmaps, maps, map_dtype, map_stride, alm, alm_stride, ws,
                       (uint64_t)HBX_STATIC_SCRATCH_BYTES, stream);
}
