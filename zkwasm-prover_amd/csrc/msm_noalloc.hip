// zk_msm_noalloc: zk_msm for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  The pipeline's scratch is a file-scope device array instead of
// a caller workspace; the call enqueues kernels only (device window
// combination, no host wait, no host or device allocation).
#include <atomic>

#include "guard.h"

#include "zkcore.h"
#include "zkinternal.h"

namespace {

// Every GPU the library runs on holds its own copy of this array, from the
// code object's load until the process ends.  In host code its name is a
// host-side stand-in: kernels get the per-device address that
// hipGetSymbolAddress returns, never the name.
__device__ __align__(256) uint8_t scratch[ZK_MSM_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::atomic<void *> g_scratch_ptr[kMaxDevices];

// The array's address on the current device, looked up on the first call on
// that device (two threads racing on that first call store the same value).
hipError_t scratch_address(void **ptr) {
    int dev = -1;
    ZK_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    void *p = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (!p) {
        ZK_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev].store(p, std::memory_order_release);
    }
    *ptr = p;
    return hipSuccess;
}

}  // namespace

extern "C" {

hipError_t zk_msm_noalloc(void *out, const void *points, const void *scalars, int n, int c, hipStream_t stream) {
    if (n <= 0 || !out || !points || !scalars || c < 0 || c == 1 || c > 17) return hipErrorInvalidValue;
    // sizes whose scratch exceeds the array are refused, never shrunk to fit
    if (zk_msm_scratch_bytes(n, c) > (size_t)ZK_MSM_NOALLOC_SCRATCH_BYTES) return hipErrorInvalidValue;
    return zk::guard([&]() -> hipError_t {
        void *ptr = nullptr;
        ZK_CHECK(scratch_address(&ptr));
        return zk::msm_one_async(out, points, scalars, (uint32_t)n, (uint32_t)c, ptr,
                                 (size_t)ZK_MSM_NOALLOC_SCRATCH_BYTES, stream);
    });
}

}  // extern "C"
