// sfg_evolve_noalloc: batched stepping that allocates nothing, for callers
// that cannot allocate device memory (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc).  The per-parcel launch constants
// and the stage buffers live in `scratch`, a file-scope device array of
// SFG_NOALLOC_BYTES that every GPU the library runs on holds until the
// process ends.  Host code uses the array only through its device address,
// taken once per device (with that device current) by hipGetSymbolAddress.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/specfab_gpu.h"
#include "sfg_noalloc.h"

namespace {

constexpr int MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[SFG_NOALLOC_BYTES / sizeof(double)];

std::mutex g_mu;
double* g_scratch[MAX_DEVICES] = {};   // device address of `scratch` per device

}  // namespace

namespace sfg {

// the device address of `scratch` on device dev, taken once with dev current;
// sfg_ctx_create calls this for its device, so that the array (loaded with
// this file's code object) is in place before any no-allocation call
int noalloc_setup(int dev, double** out) {
    if (dev < 0 || dev >= MAX_DEVICES) return set_error(SFG_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        int cur = 0;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess) e = hipSetDevice(dev);
        if (e == hipSuccess) e = hipGetSymbolAddress((void**)&g_scratch[dev], HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            g_scratch[dev] = nullptr;
            char msg[160];
            snprintf(msg, sizeof(msg), "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return set_error(SFG_ERR_HIP, msg);
        }
        (void)hipSetDevice(cur);
    }
    if (out) *out = g_scratch[dev];
    return SFG_OK;
}

}  // namespace sfg

extern "C" int sfg_evolve_noalloc(sfg_ctx* ctx, const sfg_evolve_desc* desc, void* stream) {
    if (!ctx || !desc) return sfg::set_error(SFG_ERR_ARG, "ctx/desc is NULL");
    double* ptr = nullptr;
    const int rc = sfg::noalloc_setup(sfg::ctx_device(ctx), &ptr);
    if (rc) return rc;
    return sfg::evolve_in_workspace(ctx, desc, stream, ptr, SFG_NOALLOC_BYTES);
}
