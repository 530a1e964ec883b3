// Initial states from structure tensors, DDRX rate factor, power-spectrum
// bounds (see sfg_init.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfg {

// a2 (P,3,3) row-major -> rows 0..5 of the real SoA state x (n, P), rows >= 6 zeroed
hipError_t launch_a2_to_state(int64_t P, int n, const double* a2, double* x, hipStream_t s);
// a4 (P,3,3,3,3) row-major -> rows 0..14, rows >= 15 zeroed (n >= 15)
hipError_t launch_a4_to_state(int64_t P, int n, const double* a4, double* x, hipStream_t s);
// Gamma0 (P) from strain rates eps (P,3,3), temperatures T (P) or Tconst (T null)
hipError_t launch_gamma0(int64_t P, const double* eps, const double* T, double Tconst, double A, double Q, double* G,
                         hipStream_t s);
// apply_bounds in place on the real SoA state (n, P)
hipError_t launch_bounds(int64_t P, int n, double* x, hipStream_t s);

}  // namespace sfg
