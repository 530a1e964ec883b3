// Rotation of states by Wigner-D matrices, a4 eigentensors and state
// validity (see sfg_rotate.hip).
#pragma once
#ifdef SFG_HOST_EMUL
#include "sfg_emul.h"
#else
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

namespace sfg {

constexpr int ROT_LMAX = 12;   // src/rotation.f90:37 (Lmax_wignerD)

// number of terms of Wigner's sum formula for d^l_{mn}, and the index of
// entry (l, m, n)'s first term in the table of wigner_terms (entries l = 2,
// 4, ..., m = 0..l, n = -l..l, in that order)
__host__ __device__ constexpr int wigner_count(int l, int m, int n) {
    return ((l + n < l - m) ? l + n : l - m) - ((n - m > 0) ? n - m : 0) + 1;
}
__host__ __device__ constexpr int wigner_offset(int l, int m, int n) {
    int t = 0;
    for (int ll = 2; ll <= l; ll += 2)
        for (int mm = 0; mm <= ll; mm++)
            for (int nn = -ll; nn <= ll; nn++) {
                if (ll == l && mm == m && nn == n) return t;
                t += wigner_count(ll, mm, nn);
            }
    return t;
}

struct RotateArgs {
    int64_t P;
    int L;                   // even, 2 <= L <= 12 (blocks above L are not touched)
    const double* x;         // real SoA state (nlm_len(L), P)
    double* out;             // (nlm_len(L), P); may be x
    const double* theta;     // (P) colatitude rotation
    const double* phi;       // (P) longitude rotation
    const double* coef;      // Wigner term coefficients (wigner_terms order)
};
hipError_t launch_rotate(const RotateArgs& a, hipStream_t s);
template <int LM>
__global__ void rotate_kernel(RotateArgs a);

struct A4EigArgs {
    int64_t P;
    const double* x;         // real SoA state (n, P), n >= 15 (or the (15, P) validity states)
    const double* A2;        // (6, 15) Mandel a2 map
    const double* A4F;       // (15, 15) full a4 map (REAL(4) body)
    double* Q;               // (P, 6, 3, 3) eigentensors, or null
    double* ev;              // (P, 6) eigenvalues, ascending, or null
    int* valid;              // (P) a2 and a4 eigenvalues inside (0, 1), or null
};
hipError_t launch_a4eig(const A4EigArgs& a, hipStream_t s);
__global__ void a4eig_kernel(A4EigArgs a);

}  // namespace sfg
