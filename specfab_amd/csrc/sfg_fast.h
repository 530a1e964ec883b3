// Register-resident steppers (generated per truncation L): availability and launch.
#pragma once
#include <hip/hip_runtime.h>

#include "sfg_kernels.h"

namespace sfg {
// prec: 0 = FP64, 1 = FP32 with compensated state update
bool fast_available(int L, int prec);
hipError_t launch_evolve_fast(int L, int prec, const EvolveArgs& a, hipStream_t s);
}  // namespace sfg
