// Discrete-grain path: lattice rotation of individual grain axes and the
// harmonic expansion of a weighted grain ensemble (see sfg_discrete.hip).
#pragma once
#ifdef SFG_HOST_EMUL
#include "sfg_emul.h"
#else
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

namespace sfg {

struct RiLrotArgs {
    int64_t N;              // grains
    int Nt;                 // trajectory length (Nt - 1 Euler steps)
    double dt, iota;
    const double* ri0;      // (N, 3)
    const double* D;        // (Nt, 3, 3) strain rates, row-major
    const double* W;        // (Nt, 3, 3) spins
    double* ri;             // (Nt, N, 3) trajectory, ri[0] = ri0
};
hipError_t launch_ri_lrot(const RiLrotArgs& a, hipStream_t s);
__global__ void ri_lrot_kernel(int64_t N, int Nt, double dt, double iota, const double* __restrict__ ri0,
                               const double* __restrict__ D, const double* __restrict__ W, double* __restrict__ ri);

struct RiMomentArgs {
    int64_t E, N;           // ensembles, grains per ensemble
    int order;              // 2, 4 or 6: moment tensor order
    const double* ri;       // (E, N, 3)
    const double* wi;       // (N) weights (shared by the ensembles)
    double* out;            // order 2: (E, 9) a2; 4: (E, 81) a4; 6: real state rows 0..27 of x (n, E)
    int n;                  // rows of x (order 6)
};
hipError_t launch_ri_moments(const RiMomentArgs& a, hipStream_t s);

}  // namespace sfg
