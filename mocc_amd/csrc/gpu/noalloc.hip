// mocgpu_attenuation_noalloc: the sweep kernel's attenuation for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc). mocgpu_attenuation stages its input and
// output in a buffer it allocates and frees on every call; this entry point
// stages them in a fixed device array instead and calls no hipMalloc or
// hipFree.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../../include/mocgpu.h"
#include "kernels.hpp"

namespace {

constexpr int kMax = MOCGPU_ATTENUATION_NOALLOC_MAX;

// The staged input values in [0, kMax), the results in [kMax, 2 kMax). Each
// device holds its own copy for the life of the process; the calls on one
// device share it.
static __device__ __align__(256) float scratch[2 * kMax];

// Device address of `scratch` on each device, looked up once per device with
// that device current (the array's name is a host-side stand-in in host code,
// not a device address).
constexpr int kMaxDevices = 64;
std::mutex g_mu;
float *g_ptr[kMaxDevices] = {};

hipError_t scratch_on_current_device(float **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_ptr[dev] = (float *)p;
    }
    *out = g_ptr[dev];
    return hipSuccess;
}

}  // namespace

extern "C" int32_t mocgpu_attenuation_noalloc(const float *tau, float *out, int32_t n)
{
    if (n <= 0) return 0;
    if (!tau || !out) return MOCGPU_ERR_ARG;
    if (n > kMax) return MOCGPU_ERR_CAPACITY;
    float *d = nullptr;
    if (scratch_on_current_device(&d) != hipSuccess) return MOCGPU_ERR_DEVICE;
    // every element the kernel reads is copied in first, and it writes every
    // output element it returns: nothing of an earlier call is read
    hipError_t e = hipMemcpy(d, tau, sizeof(float) * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        mocc_amd::gpu::launch_attenuation(d, d + kMax, n, nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, d + kMax, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost);
    return e == hipSuccess ? 0 : MOCGPU_ERR_DEVICE;
}
