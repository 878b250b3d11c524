// Error reporting, version and the workspace pool for the C ABI declared in
// include/pcvl_hip.h.
#include "common.h"
#include "../../include/pcvl_hip.h"

#include <mutex>

namespace pcvl {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

// One stream-ordered pool per device, created on first use.  Up to its
// release threshold (512 MiB), memory freed into it stays mapped for the next
// call instead of going back to the device at each synchronisation, which
// made the first launch after every synchronisation map its workspace again
// (0.3-0.5 ms on a 65,536-proposal MIS chunk).  Beyond the threshold, memory
// is returned at synchronisation as the default pool does, so one large
// workspace (a mask selection over 10^9 states) is not held for good.
hipError_t ws_alloc(void **p, size_t bytes, hipStream_t stream)
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static hipMemPool_t pools[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    hipMemPool_t pool = nullptr;
    {
        std::lock_guard<std::mutex> lock(mu);
        if (!pools[dev]) {
            hipMemPoolProps props = {};
            props.allocType = hipMemAllocationTypePinned;
            props.handleTypes = hipMemHandleTypeNone;
            props.location.type = hipMemLocationTypeDevice;
            props.location.id = dev;
            hipMemPool_t created = nullptr;
            e = hipMemPoolCreate(&created, &props);
            if (e != hipSuccess) return e;
            uint64_t threshold = 512ull << 20;
            e = hipMemPoolSetAttribute(created, hipMemPoolAttrReleaseThreshold, &threshold);
            if (e != hipSuccess) {
                (void)hipMemPoolDestroy(created);
                return e;
            }
            pools[dev] = created;
        }
        pool = pools[dev];
    }
    return hipMallocFromPoolAsync(p, bytes, pool, stream);
}

} // namespace pcvl

extern "C" const char *pcvl_last_error(void) { return pcvl::g_last_error.c_str(); }

extern "C" int pcvl_abi_version(void) { return 100; }
