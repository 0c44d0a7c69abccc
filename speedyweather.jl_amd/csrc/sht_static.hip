// Allocation-free transform entry point: the Fourier scratch of the call lives in a
// file-scope device array instead of the per-(plan, stream) buffers the regular entry
// points allocate on a stream's first call.
//
// Every GPU the library runs on holds its own copy of `scratch` from the time the library's
// device code is loaded there until the process ends.  Its address on a device comes from
// hipGetSymbolAddress, taken once per device when the first plan is created on that device
// (sht_plan_create runs static_scratch_setup with the device current); host code passes
// only that address to kernels, never the array's name.
#include <mutex>
#include <string>

#include "sht_internal.h"

namespace {

constexpr size_t kStaticScratchBytes = size_t(SHT_STATIC_SCRATCH_BYTES);
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

int sht::static_scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_last_error("device ordinal " + std::to_string(device) + " beyond the static scratch table");
        return SHT_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_scratch_ptr[device]) return SHT_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return SHT_ERR_HIP;
    }
    g_scratch_ptr[device] = ptr;
    return SHT_OK;
}

extern "C" int sht_transform_static(sht_plan *plan, int32_t direction, const void *in, int64_t ld_in, void *out,
                                    int64_t ld_out, int32_t K, int32_t unscale_coslat, void *stream) {
    sht::TraceRange trace_range("sht_transform_static");
    if (!plan) {
        sht::set_last_error("null plan");
        return SHT_ERR_ARG;
    }
    if (K < 1) {
        sht::set_last_error("DimensionMismatch: K=" + std::to_string(K) + " columns");
        return SHT_ERR_DIM;
    }
    const size_t need = sht::plan_scratch_bytes(plan, K);
    if (need > kStaticScratchBytes) {
        sht::set_last_error("sht_transform_static: K=" + std::to_string(K) + " needs " + std::to_string(need) +
                            " bytes of scratch, more than the static " + std::to_string(kStaticScratchBytes) +
                            " (use sht_synthesis / sht_analysis)");
        return SHT_ERR_DIM;
    }
    const int dev = sht::plan_device(plan);
    void *ptr = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        if (dev >= 0 && dev < kMaxDevices) ptr = g_scratch_ptr[dev];
    }
    if (!ptr) {
        sht::set_last_error("static scratch of device " + std::to_string(dev) + " is not set up");
        return SHT_ERR_HIP;
    }
    return sht::transform_with_scratch(plan, direction, in, ld_in, out, ld_out, K, unscale_coslat, ptr, stream);
}
