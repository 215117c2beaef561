// Allocation-free FP64 residual (ice_residual_noalloc).
//
// ice_residual keeps its per-call buffers (element face traces, face fluxes,
// DDGIC lifting vectors, wave-speed partials) in memory the discretization
// allocated with hipMalloc.  Callers that must not allocate at all — code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc — use this entry point instead: the same kernels run with those
// buffers carved from one fixed-size array in device memory, `scratch`,
// which every GPU the library runs on holds until the process ends.
//
// The host may not use the array's name as a device address (it is a
// host-side stand-in): the device address is looked up once per device with
// hipGetSymbolAddress on the first call made with that device current, and
// only that pointer is handed to the kernels.  Every call on one GPU shares
// that GPU's array, so calls that overlap on one GPU (two streams, two
// threads) race; calls ordered on one stream do not.  The kernels write
// every scratch entry they later read within the same call, so nothing in
// the array needs clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "engine.h"

namespace {

constexpr size_t kScratchBytes = (size_t)ICE_NOALLOC_SCRATCH_MB << 20;
constexpr size_t kScratchDoubles = kScratchBytes / sizeof(double);
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_scratch_mtx;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's copy of `scratch` (looked up once per device).
int scratch_of_current_device(double** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        ice::set_error("ice_residual_noalloc: no current HIP device (or device id beyond 63)");
        return ICE_ERR_HIP;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mtx);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            ice::set_error(std::string("ice_residual_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                           hipGetErrorString(e));
            return ICE_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return ICE_OK;
}

}  // namespace

extern "C" int ice_noalloc_scratch_bytes(const ice_disc* disc, int64_t* need_bytes, int64_t* max_bytes) {
    if (max_bytes) *max_bytes = (int64_t)kScratchBytes;
    if (!disc) {
        if (need_bytes) *need_bytes = 0;
        return need_bytes ? (ice::set_error("null discretization"), ICE_ERR_ARG) : ICE_OK;
    }
    size_t need = 0;
    ice::disc_residual_in(const_cast<ice_disc*>(disc), nullptr, nullptr, nullptr, nullptr, 0, &need);
    if (need_bytes) *need_bytes = (int64_t)(need * sizeof(double));
    return ICE_OK;
}

extern "C" int ice_residual_noalloc(ice_disc* disc, const double* u_dev, double* res_dev, void* stream) {
    if (!disc || !u_dev || !res_dev) {
        ice::set_error("null argument");
        return ICE_ERR_ARG;
    }
    size_t need = 0;
    ice::disc_residual_in(disc, nullptr, nullptr, nullptr, nullptr, 0, &need);
    if (need > kScratchDoubles) {
        ice::set_error("ice_residual_noalloc: the discretization needs " + std::to_string(need * sizeof(double)) +
                       " bytes of scratch, more than the fixed " + std::to_string(kScratchBytes) +
                       " (ICE_NOALLOC_SCRATCH_MB); use ice_residual");
        return ICE_ERR_ARG;
    }
    double* ptr = nullptr;
    if (const int rc = scratch_of_current_device(&ptr); rc != ICE_OK) return rc;
    const int rc = ice::disc_residual_in(disc, u_dev, res_dev, (hipStream_t)stream, ptr, kScratchDoubles, &need);
    if (rc != 0) {
        ice::set_error(std::string("ice_residual_noalloc: residual launch failed: ") +
                       hipGetErrorString((hipError_t)rc));
        return ICE_ERR_HIP;
    }
    return ICE_OK;
}
