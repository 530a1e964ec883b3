// Reduced real form of the expansion (rnlm) and of the fabric operators
// (Mrr, Mri, Mir, Mii), batched over mesh nodes (see sfg_reduced.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfg {

// number of m >= 0 coefficients with even l <= L: (L+2)^2/4
inline int rnlm_len(int L) { return (L + 2) * (L + 2) / 4; }

// M (P, n, n) complex row-major -> Mrr, Mri, Mir, Mii (P, rn, rn) real row-major
hipError_t launch_reduce_M(int64_t P, int L, const double* M, double* Mrr, double* Mri, double* Mir, double* Mii,
                           hipStream_t s);
// nlm (P, n) complex -> rnlm (P, rn) complex
hipError_t launch_nlm_to_rnlm(int64_t P, int L, const double* nlm, double* rnlm, hipStream_t s);
// rnlm (P, rn) complex -> nlm (P, n) complex
hipError_t launch_rnlm_to_nlm(int64_t P, int L, const double* rnlm, double* nlm, hipStream_t s);

}  // namespace sfg
