// Kernel argument blocks and launchers (host-visible declarations).
#pragma once
#ifdef SFG_HOST_EMUL
#include "sfg_emul.h"   // CPU build of the generated kernels (tests only)
#else
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

// Constant address space for read-only coupling tables: loads through such a
// pointer with a uniform address become scalar (SGPR) loads.
#ifdef SFG_HOST_EMUL
#define SFG_CONST_AS
#else
#define SFG_CONST_AS __attribute__((address_space(4)))
#endif

namespace sfg {

constexpr int NPAR_K = 23;   // 8 LROT + 15 DDRX parameters

struct ParamArgs {
    int64_t P;
    int nblk;          // number of l blocks (L/2 + 1)
    int nq;            // min(n, 15): coefficients entering <D>
    int procs;
    double iota, zeta, reg_nu;
    const double* ugrad;   // (P,3,3)
    const double* tau;     // (P,3,3) or null
    const double* gam0;    // (P) or null
    const double* lam;     // (P) or null
    const double* A2;      // (6,15)
    const double* A4;      // (36,15)
    const double* reg_shape;  // (nblk)
    double* theta;         // (NPAR_K, P)
    double* omega;         // (15, P)
    double* sl;            // (nblk, P)
};

struct EvolveArgs {
    int64_t P;
    int n, nq, scheme, nsteps, use_ddrx;
    const int* steps;
    double dt;
    const int2* rows;      // (n) group ranges
    const int* row_blk;    // (n) l/2
    const int* grp_param;
    const int2* grp_rng;
    const int* tcol;
    const double* tval;
    const float* tval_f;   // the same coupling values rounded to FP32 (FP32 stepping)
    const double* theta;
    const double* omega;
    const double* sl;
    double* X;
    double* YS;
    double* XA;
    double* XB;
    int* status;
    int64_t zero;          // always 0 (an opaque stride for the generated kernels)
};

struct DenseArgs {
    int64_t P;
    int n, mode;
    double iota, zeta;
    const double* A;       // eps or tau (P,3,3)
    const double* B;       // omg (P,3,3)
    const int* row_ptr;    // (NMAX+1)
    const int4* ent;       // (t, j, k, 0)
    const double* val;
    double* M;             // (P,n,n) complex
};

hipError_t launch_params(const ParamArgs& a, hipStream_t s);
hipError_t launch_evolve_generic(const EvolveArgs& a, hipStream_t s);
hipError_t launch_nlm_to_state(int64_t P, int n, const int* r2c, const int* rpart, const double* nlm, double* x,
                               hipStream_t s);
hipError_t launch_state_to_nlm(int64_t P, int n, const int* cre, const int* cim, const int* csgn, const int* cneg,
                               const double* x, double* nlm, hipStream_t s);
hipError_t launch_a2(int64_t P, const double* x, double* a2, hipStream_t s);
hipError_t launch_a4mandel(int64_t P, int nq, const double* A4, const double* x, double* a4v, hipStream_t s);
hipError_t launch_dense_op(const DenseArgs& a, hipStream_t s);
hipError_t launch_evD2(int64_t P, int nq, const double* A2, const double* A4, const double* x, const double* tau,
                       double* out, hipStream_t s);
hipError_t launch_diag_op(int64_t P, int n, int mode, const int* row_l, const double* reg_shape, double reg_nu,
                          const double* D, double* M, hipStream_t s);

}  // namespace sfg
