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
    double* Qn;           // (P, ndir, 9) the acoustic tensors (row-major), or null
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
    const double* R;      // Vi: (ndir, ROT4_STRIDE)
    const double* frames; // Eij: (P, 3, 3) rows e1, e2, e3, or null (Cartesian)
    double ci[6];         // Eij: 4/3 lami(1:3), 2 lami(4:6) (rheo_params_orthotropic)
    double q6iso[36];     // Eij: the isotropic fabric's Sachs tensor, pair form (set by the launch)
    double dqiso;         //      and its (3,2,1,2) excess
    double lam[6], mu[3], rho;   // Vi: grain Lame parameters, density
    double* Eij;          // (P, 6), or (P, nstates) for Evw
    double* Vi;           // (P, ndir, 3): qS1, qS2, qP
    const double* states; // Evw: (nstates, 15) = v(3), w(3), tau(9) row-major (tau symmetric), device
    int nstates;
};
// Orthotropic grains given as discrete axes (src/moments.f90:386-417): E
// ensembles of N grains, each axis array (E, N, 3).  The moment pass sums,
// per ensemble and axis m, the 57 grain averages a2_m (6 pairs), a4_m (15
// sorted entries) and J_m = <q q q' q'> (6 x 6 pairs; m = 0: <n n v v>,
// 1: <b b v v>, 2: <b b n n>), in G partial sums per ensemble.
constexpr int ORTHO_NMOM = 57;
struct OrthoDiscArgs {
    int64_t E, N;
    int G;                  // workgroups per ensemble (grain chunks)
    const double* b;
    const double* n;
    const double* v;
    double* part;           // (E, G, 3, 57) partial sums
    double* mom;            // (E, 3, 57) grain averages
};
hipError_t launch_ortho_moments(const OrthoDiscArgs& a, hipStream_t s);
// consumers of the averages: Eij (frames as for OrthoArgs, isotropic
// denominator of the continuum isotropic fabric) and Vi along unit
// directions kdir (ndir, 3)
struct OrthoMomArgs {
    int64_t E;
    const double* mom;
    const double* frames;   // (E, 3, 3) or null (Cartesian)
    double ci[6];
    double q6iso[36];
    double dqiso;
    double* Eij;            // (E, 6)
    int ndir;
    const double* kdir;     // (ndir, 3)
    double lam[6], mu[3], rho;
    double* Vi;             // (E, ndir, 3)
};
hipError_t launch_eij_ortho_mom(const OrthoMomArgs& a, hipStream_t s);
hipError_t launch_vi_ortho_mom(const OrthoMomArgs& a, hipStream_t s);
// Q6 (pair form) and the (3,2,1,2) excess of the isotropic fabric's Sachs tensor
void ortho_q6_iso(const double ci[6], double Q6[36], double* dq);
hipError_t launch_eij_ortho(const OrthoArgs& a, hipStream_t s);
hipError_t launch_vi_ortho(const OrthoArgs& a, hipStream_t s);
hipError_t launch_evw_ortho(const OrthoArgs& a, hipStream_t s);

hipError_t launch_vi_elastic(const ViElasticArgs& a, hipStream_t s);
hipError_t launch_vi_em(const ViEMArgs& a, hipStream_t s);
hipError_t launch_a4full(const A4FullArgs& a, hipStream_t s);

}  // namespace sfg
