// Homogenization kernel arguments and launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfg {

struct EijArgs {
    int64_t P;
    int nq;
    int frame;            // 0 = a2 eigenframe, 1 = given frames, 2 = Cartesian (x, y, z)
    int n_grain;          // 1 or -3
    double Ecc, Eca, alpha;
    const double* x;      // real SoA state (n, P)
    const double* A2;     // (6, 15)
    const double* A4;     // (36, 15)
    const double* frames; // (P, 3, 3) rows e1, e2, e3 (frame == 1)
    double* Eij;          // (P, 6)
    double* lami;         // (P, 3) or null
    double* ei;           // (P, 3, 3) or null
    int* status;          // (P) or null
    // optional work list: items t < nidx evaluate state column idx[t] and
    // write Eij/lami/ei at row t (status stays at the parcel's own row);
    // with nidx_dev the item count is min(*nidx_dev, nidx), read on the
    // device (nidx is then the list's capacity), so no host round trip
    const int* idx;
    int64_t nidx;
    const int* nidx_dev;
};

hipError_t launch_eij(const EijArgs& a, hipStream_t s);

}  // namespace sfg
