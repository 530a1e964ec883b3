// Kernel argument blocks and launchers (host-visible declarations).
#pragma once
#ifdef SFG_HOST_EMUL
#include "sfg_emul.h"   // CPU build of the generated kernels (tests only)
#else
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

namespace sfg {

// Raw buffer access for the generated steppers: a wave-uniform resource
// (base, byte size) in SGPRs, one shared VGPR lane offset and a scalar row
// offset, so that SoA rows cost no per-access vector address arithmetic.
// Reads beyond the resource's size return 0.
#ifndef SFG_HOST_EMUL
typedef __amdgpu_buffer_rsrc_t sfg_rsrc;
__device__ __forceinline__ sfg_rsrc make_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ double bld64(sfg_rsrc r, uint32_t vo, uint32_t so) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 0));
}
__device__ __forceinline__ void bst64(sfg_rsrc r, uint32_t vo, uint32_t so, double v) {
    typedef int i32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(i32x2, v), r, vo, so, 0);
}
__device__ __forceinline__ float bld32(sfg_rsrc r, uint32_t vo, uint32_t so) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
}
__device__ __forceinline__ void bst32(sfg_rsrc r, uint32_t vo, uint32_t so, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, vo, so, 0);
}
#else
struct sfg_rsrc { char* base; uint32_t bytes; };
inline sfg_rsrc make_rsrc(const void* p, uint32_t bytes) { return sfg_rsrc{(char*)p, bytes}; }
template <class T>
inline T bld(sfg_rsrc r, uint32_t vo, uint32_t so) {
    uint64_t o = (uint64_t)vo + so;
    T v{};
    if (o + sizeof(T) <= r.bytes) __builtin_memcpy(&v, r.base + o, sizeof(T));
    return v;
}
template <class T>
inline void bst(sfg_rsrc r, uint32_t vo, uint32_t so, T v) {
    uint64_t o = (uint64_t)vo + so;
    if (o + sizeof(T) <= r.bytes) __builtin_memcpy(r.base + o, &v, sizeof(T));
}
inline double bld64(sfg_rsrc r, uint32_t vo, uint32_t so) { return bld<double>(r, vo, so); }
inline void bst64(sfg_rsrc r, uint32_t vo, uint32_t so, double v) { bst<double>(r, vo, so, v); }
inline float bld32(sfg_rsrc r, uint32_t vo, uint32_t so) { return bld<float>(r, vo, so); }
inline void bst32(sfg_rsrc r, uint32_t vo, uint32_t so, float v) { bst<float>(r, vo, so, v); }
#endif

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
    double reg_shape[11];  // (nblk <= 11): (l(l+1) / (L(L+1)))^expo per l block
    double* theta;         // (NPAR_K, P)
    double* omega;         // (15, P)
    double* sl;            // (nblk, P)
};

struct EvolveArgs {
    int64_t P;
    int n, nq, scheme, nsteps, use_ddrx;
    // per-parcel step counts (or null: nsteps for all).  steps_dec = 1: steps
    // holds the remaining count; a launch takes min(remaining, nsteps) steps
    // and stores the new remaining count.
    int* steps;
    int steps_dec;
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
    const int4* ent;       // (t, j, k, 0), sorted by (row, j, t, k)
    const double* val;
    const int2* seg;       // (segw, segw): [begin, end) of the entries of (row, j)
    int segw;
    double* M;             // (P,n,n) complex
};

hipError_t launch_params(const ParamArgs& a, hipStream_t s);
// The reduced blocks as a linear map of each node's coupling weights: with
// Z = sg sum_e val_e w[t_e][k_e] (sg = -1 LROT, +1 DDRX) every block entry
// is sum_c C[c] Re w_c or Im w_c, C = sg (A+ +- s A-) over the K weight
// components c (LROT: t * 6 + k, K = 24; DDRX: k, K = 15).  Cp, Cm (K,
// rn^2); W (P, 48) the weights' real parts then imaginary parts (stride 24).
// Most (c, entry) coefficients are structural zeros (M_LROT at L = 12: 2,391
// of 57,624 nonzero, at most 10 per entry for every L; M_DDRX_src at most 5),
// so the map is stored as an ELL table: entry r's nonzero terms in
// increasing c fill slots s < S (slot-major, ci / cp / cm[s * rn2 + r];
// unused slots have c = 0 and zero coefficients), and nt[r / 64] is the
// largest term count of the 64 entries of a wave.
constexpr int RMAP_W = 48;
constexpr int RMAP_NODES = 16;   // nodes per workgroup chunk (their weights staged in LDS)
struct RMap {
    int rn2, S;                  // S: slots per entry (10 LROT, 5 DDRX)
    const int* ci;               // (S, rn2) weight component of each slot
    const double* cp;            // (S, rn2) coefficient of Re w (Mrr) and Im w (Mir)
    const double* cm;            // (S, rn2) coefficient of Re w (Mii) and -Im w (Mri)
    const int* nt;               // (ceil(rn2 / 64)) trips per wave
};
hipError_t launch_dense_weights(const DenseArgs& a, double* W, hipStream_t s);
hipError_t launch_reduced_map(int64_t P, const RMap& m, const double* W, double* Mrr, double* Mri, double* Mir,
                              double* Mii, hipStream_t s);
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
