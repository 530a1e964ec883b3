// Wave-velocity and full-a4 kernel arguments and launchers.
#pragma once
#ifdef SFG_HOST_EMUL
#include "sfg_emul.h"
#else
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

namespace sfg {

constexpr int ROT4_STRIDE = 25 + 81;   // per direction: R2 (5x5) then R4 (9x9), row-major

struct ViElasticArgs {
    int64_t P;
    int ndir, nq;
    const double* x;      // real SoA state (n, P)
    const double* A2;     // (6, 15) Mandel a2 map
    const double* A4;     // (36, 15) Mandel a4 map (REAL(8) body)
    const double* A4F;    // (15, 15) full a4 map (REAL(4) body), index-sorted entries
    const double* R;      // (ndir, ROT4_STRIDE)
    double alpha, rho;
    double kf[5];         // elas_fwdparams (Reuss)
    double kr[5];         // elas_revparams (Voigt)
    double* Vi;           // (P, ndir, 3): qS1, qS2, qP
    int* status;          // (P) or null
    // optional work list: items t < nidx evaluate state column idx[t] and
    // write Vi at row t; with nidx_dev the item count is min(*nidx_dev,
    // nidx), read on the device (nidx is then the list's capacity)
    const int* idx;
    int64_t nidx;
    const int* nidx_dev;
};

struct ViEMArgs {
    int64_t P;
    int ndir;
    const double* x;
    const double* A2;
    const double* R;      // (ndir, ROT4_STRIDE); only R2 is read
    double eps_m, eps_t, mu;
    double* Vi;           // (P, ndir, 2): S1 (fast), S2 (slow)
};

struct A4FullArgs {
    int64_t P;
    int nq;
    const double* x;
    const double* A4F;
    double* a4;           // (P, 3, 3, 3, 3) row-major
};

// Orthotropic grains (olivine): three distributions q1, q2, q3 of the
// slip-system axes (b, n, v) per parcel, or q1, q2 with v = b x n derived
// (src/moments.f90:357-384).  Eij (Sachs, src/enhancementfactors.f90:134-188)
// or elastic Vi (Voigt, src/wavepropagation.f90:41-68,98-115).
struct OrthoArgs {
    int64_t P;
    int nq;               // rows of the l <= 4 block present (15 for L >= 4)
    int ndir;             // Vi: directions
    const double* x1;     // (n, P) real states of q1, q2, q3 (x3 null: v = b x n)
    const double* x2;
    const double* x3;
    const double* A2;     // (6, 15) Mandel a2 map
    const double* A4F;    // (15, 15) full a4 map (REAL(4) body), index-sorted entries
    const double* R;      // Vi: (ndir, ROT4_STRIDE)
    const double* frames; // Eij: (P, 3, 3) rows e1, e2, e3, or null (Cartesian)
    double ci[6];         // Eij: 4/3 lami(1:3), 2 lami(4:6) (rheo_params_orthotropic)
    const double* Qiso;   // Eij: the isotropic fabric's Sachs tensor Q (81, row-major lkij; device)
    double lam[6], mu[3], rho;   // Vi: grain Lame parameters, density
    double* Eij;          // (P, 6)
    double* Vi;           // (P, ndir, 3): qS1, qS2, qP
};
// Eij: Qiso is device scratch of 81 doubles for the isotropic fabric's Q
hipError_t launch_eij_ortho(const OrthoArgs& a, double* Qiso, hipStream_t s);
hipError_t launch_vi_ortho(const OrthoArgs& a, hipStream_t s);

hipError_t launch_vi_elastic(const ViElasticArgs& a, hipStream_t s);
hipError_t launch_vi_em(const ViEMArgs& a, hipStream_t s);
hipError_t launch_a4full(const A4FullArgs& a, hipStream_t s);

}  // namespace sfg
