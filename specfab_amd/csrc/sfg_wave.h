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

hipError_t launch_vi_elastic(const ViElasticArgs& a, hipStream_t s);
hipError_t launch_vi_em(const ViEMArgs& a, hipStream_t s);
hipError_t launch_a4full(const A4FullArgs& a, hipStream_t s);

}  // namespace sfg
