// Minimal host stand-ins for the HIP device vocabulary used by the generated
// stepping kernels, so that tests can compile and run a generated kernel on
// the CPU lane by lane (one lane's program order at a time) and compare it
// with the oracle.  Used only by the CPU emulation test build (SFG_HOST_EMUL).
#pragma once
#include <math.h>
#include <cmath>
#include <cstdint>

struct int2 { int x, y; };
struct int4 { int x, y, z, w; };
struct sfg_dim3 { unsigned x, y, z; };
extern thread_local sfg_dim3 threadIdx, blockIdx, blockDim, gridDim;
typedef int hipError_t;
typedef void* hipStream_t;
#define __global__
#define __device__
#define __host__
#define __launch_bounds__(...)
#define __builtin_amdgcn_sched_barrier(m) ((void)0)
// workgroup barrier of the emulated block (the driver runs a block's threads
// concurrently; see tests/emul_driver.cpp); a no-op where the driver runs
// threads one after another
void sfg_emul_syncthreads();
#define __syncthreads() sfg_emul_syncthreads()
#define __shared__ static
#define __forceinline__ inline
#define __constant__
using std::fma;
using std::isfinite;
inline int atomicOr(int* p, int v) {
    int o = *p;
    *p |= v;
    return o;
}
