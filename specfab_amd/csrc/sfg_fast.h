// Register-resident steppers (generated per truncation L and variant):
// availability and launch.
#pragma once
#include <hip/hip_runtime.h>

#include "sfg_kernels.h"

namespace sfg {
// prec: 0 = FP64, 1 = FP32 with compensated state update.  Number of
// generated variants of (L, prec) (0: no register-resident stepper).
int fast_variants(int L, int prec);
hipError_t launch_evolve_fast(int L, int prec, int variant, const EvolveArgs& a, hipStream_t s);
// the variant's __global__ function (for graph kernel nodes), its threads per
// workgroup and the parcels one workgroup steps (fewer than the threads when
// several waves share a parcel group): grid = ceil(P / parcels)
const void* fast_kernel(int L, int prec, int variant, int* block, int* parcels);
// stage buffers of EvolveArgs that a variant uses (the others may be null):
// the workspace holds only those, each n x P doubles
enum { BUF_YS = 1, BUF_XA = 2, BUF_XB = 4 };
int fast_buffers(int L, int prec, int variant);
// the table-driven stepper's __global__ function and its block size
const void* generic_kernel();
constexpr int GENERIC_BLOCK = 64;
inline bool fast_available(int L, int prec) { return fast_variants(L, prec) > 0; }
}  // namespace sfg
