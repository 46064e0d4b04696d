// Allocation-free directory runs (fw_run_dir_noalloc): every device array of the run lives in
// one static device array per GPU, so the call makes no device allocation.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/fullwave_hip.h"
#include "fw_internal.h"
#include "fw_sim.h"

// The device memory of allocation-free runs. Each GPU the library runs on holds its own copy
// from the module's load on that GPU until the process ends. In host code the name is only a
// stand-in for the symbol, never a device address: the address on each device comes from
// hipGetSymbolAddress (scratch_on).
static __device__ __align__(256) char scratch[FW_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_lock[kMaxDevices];    // one allocation-free run per device at a time
void* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device, once resolved

// The address of `scratch` on `device`, resolved once per device with that device current.
// Called with g_lock[device] held.
int scratch_on(int device, void** out) {
    if (!g_scratch[device]) {
        int prev = -1;
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        hipError_t e = hipSetDevice(device);
        void* p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev >= 0) (void)hipSetDevice(prev);
        if (e != hipSuccess || !p)
            return fw::set_error(FW_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                                 hipGetErrorString(e));
        g_scratch[device] = p;
    }
    *out = g_scratch[device];
    return FW_OK;
}

}  // namespace

extern "C" int fw_run_dir_noalloc(const char* sim_dir, int device) {
    fw_desc d{};
    int rc = fw::read_dir_desc(sim_dir, &d);
    if (rc != FW_OK) return rc;
    // sizes are checked before any device work
    const size_t need = fw::fixed_footprint(d);
    if (need > (size_t)FW_NOALLOC_SCRATCH_BYTES)
        return fw::set_error(FW_ERR_UNSUPPORTED,
                             "this run needs %zu bytes of device memory; the allocation-free entry point holds at "
                             "most %zu (use fw_run_dir)",
                             need, (size_t)FW_NOALLOC_SCRATCH_BYTES);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev || device >= kMaxDevices)
        return fw::set_error(FW_ERR_DEVICE, "device %d not available (%d devices)", device, ndev);
    std::lock_guard<std::mutex> lock(g_lock[device]);
    void* base = nullptr;
    if ((rc = scratch_on(device, &base)) != FW_OK) return rc;
    return fw::run_dir_impl(sim_dir, device, base, (size_t)FW_NOALLOC_SCRATCH_BYTES);
}
