/*
 * hpm_noalloc.hip — the learned prior's network forward without any device
 * allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).
 *
 * hpm_lprior_forward allocates its activation area per call. This variant
 * runs the same kernels over `scratch`, a fixed array in the code object
 * that every GPU the library runs on holds until the process ends. Host
 * code never passes the array's name to a kernel (on the host it is only a
 * stand-in for the symbol): the device address comes from
 * hipGetSymbolAddress, looked up once per device with that device current,
 * and only that address is handed to the kernels.
 */
#include <hip/hip_runtime.h>

#include <mutex>

#include "hpm_internal.h"
#include "hpm_unet.h"

/* 64 MiB: the activations of one forward pass take about 552 bytes per
 * pixel (BF16 maps of the four levels), so up to about 120k pixels, e.g.
 * 400 x 300 or a 256 x 256 training crop. */
#define HPM_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

static __device__ __align__(256) unsigned char scratch[HPM_NOALLOC_SCRATCH_BYTES];

namespace {

const int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

/* the current device's copy of `scratch` (looked up on first use) */
hipError_t scratch_address(int device, void** out)
{
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[device] == nullptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            return e;
        }
        g_scratch[device] = p;
    }
    *out = g_scratch[device];
    return hipSuccess;
}

} // namespace

extern "C" size_t hpm_lprior_noalloc_capacity(void)
{
    return HPM_NOALLOC_SCRATCH_BYTES;
}

extern "C" int hpm_lprior_forward_noalloc(hpm_lprior* lp, const void* x, int height, int width, float* head,
                                          void* stream)
{
    if (!lp || !x || !head || height < 1 || width < 1) {
        return HPM_EINVAL;
    }
    if (hpm_lprior_activation_bytes(width, height) > HPM_NOALLOC_SCRATCH_BYTES) {
        return HPM_ETOOBIG;
    }
    const int device = hpm_lprior_device(lp);
    if (device < 0 || device >= kMaxDevices) {
        return HPM_EINVAL;
    }
    hipError_t e = hipSetDevice(device);
    void* ws = nullptr;
    if (e == hipSuccess) {
        e = scratch_address(device, &ws);
    }
    if (e == hipSuccess) {
        e = hpm_lprior_forward_ws(lp, x, height, width, head, ws, (hipStream_t)stream);
    }
    return (int)e;
}
