// av1pf_static.hip - vaco_av1pf_frame_device_static: the per-frame stage
// without allocating anything, for callers that cannot allocate device
// memory.
//
// Every device buffer the call needs (the status word, the staged side
// information and work descriptors, the CDEF and upscaled intermediate
// images) lives in `scratch`, a file-scope device array of
// VACO_AV1PF_STATIC_SCRATCH_BYTES. Each GPU holds its own copy from the
// library's first use on it until the process ends. In host code the
// array's name is only a stand-in for the symbol, so its device address is
// looked up once per device (hipGetSymbolAddress, with that device current)
// and only that address is passed on: to the kernels, to the copies.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/vaco_av1pf.h"
#include "av1pf_internal.h"

namespace {

constexpr size_t kScratchBytes = VACO_AV1PF_STATIC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

// One GPU's scratch address, and the lock that keeps two of this process's
// calls on that GPU from using the array at once.
struct StaticDevice {
    std::mutex mu;
    uint8_t *ptr = nullptr;
};
StaticDevice g_devices[kMaxDevices];

}  // namespace

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

extern "C" vaco_av1pf_status vaco_av1pf_frame_device_static(const vaco_av1pf_frame_desc *desc,
                                                            const vaco_av1pf_picture *in,
                                                            vaco_av1pf_picture *out, void *stream,
                                                            const char **msg)
{
    const char *unused = "";
    if (!msg)
        msg = &unused;
    *msg = "";
    if (!desc || !in || !out) {
        *msg = "null argument";
        return VACO_AV1PF_INVALID_ARGUMENT;
    }
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        *msg = e != hipSuccess ? hipGetErrorString(e) : "device index out of range";
        return VACO_AV1PF_DEVICE_ERROR;
    }
    StaticDevice &sd = g_devices[dev];
    std::lock_guard<std::mutex> lock(sd.mu);
    if (!sd.ptr) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *msg = e != hipSuccess ? hipGetErrorString(e) : "scratch symbol has no address";
            return VACO_AV1PF_DEVICE_ERROR;
        }
        sd.ptr = static_cast<uint8_t *>(p);
    }
    return av1pf_frame_static(desc, in, out, static_cast<hipStream_t>(stream), sd.ptr, kScratchBytes, msg);
}
