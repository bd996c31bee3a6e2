/*
 * gm_noalloc.hip -- the library's fixed device array for the entry point that
 * allocates nothing (score_gm_get_status_noalloc).
 *
 * The status vector of the front (two uint32 columns, at most
 * SCORE_GM_STATUS_NOALLOC_MAX entries each, 32 MiB in all) is staged here
 * instead of in a buffer allocated per call.  Every GPU the library runs on
 * holds its own copy of the array from the first use until the process ends.
 * In host code the array's name is a host-side stand-in, not a device address:
 * its address on each device is looked up once, with that device current, and
 * only that address is handed to kernels and copies.  Calls on one GPU share
 * its array, so they are serialised per device.
 */
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/score_gm.h"
#include "gm_internal.h"

static __device__ __align__(256) uint32_t scratch[2u * SCORE_GM_STATUS_NOALLOC_MAX];

namespace {

constexpr int kMaxDevices = 64;

struct DeviceScratch {
    std::mutex use;          /* held by a call for as long as it uses the array */
    std::once_flag looked_up;
    hipError_t err = hipSuccess;
    void *ptr = nullptr;     /* the array's address on this device */
};

DeviceScratch g_scratch[kMaxDevices];

}  // namespace

hipError_t gm_scratch_acquire(int device, uint32_t **ptr, std::unique_lock<std::mutex> &lock) {
    *ptr = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    DeviceScratch &ds = g_scratch[device];
    std::unique_lock<std::mutex> lk(ds.use);
    std::call_once(ds.looked_up, [&] {
        void *p = nullptr;
        ds.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (ds.err == hipSuccess && !p) ds.err = hipErrorInvalidSymbol;
        if (ds.err == hipSuccess) ds.ptr = p;
        else (void)hipGetLastError();
    });
    if (ds.err != hipSuccess) return ds.err;
    *ptr = static_cast<uint32_t *>(ds.ptr);
    lock = std::move(lk);
    return hipSuccess;
}
