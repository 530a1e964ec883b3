// Surrogate homogenization kernels (see sfg_surrogate.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfg {

// per-parcel flags: outside the a2-eigenvalue envelope, drawn for the audit,
// over the error bound of the learned Taylor term
enum { SUR_FALLBACK = 1, SUR_AUDIT = 2, SUR_BOUND = 4 };

struct SurrogateFeatArgs {
    int64_t P;
    const double* x;       // real SoA state (n, P), n >= 15
    const double* ugrad;   // (P, 3, 3)
    const double* A2;      // (6, 15) Mandel a2 map
    float* feat;           // (P, 36)
    int* flags;            // (P) or null: SUR_FALLBACK / SUR_AUDIT / SUR_BOUND
    double env_lo[3], env_hi[3];
    double audit_frac;
    uint64_t seed;
};

// Eij = (1 - alpha) Sachs + alpha Taylor (src/enhancementfactors.f90:23-69)
// in the Cartesian frame: the Sachs term in closed form from a2 and nine +
// three entries of the Mandel a4, the Taylor term from the network (its six
// Eij outputs), and the per-parcel bound alpha T dT_q / E_q <= tol on the
// relative error that the network's calibrated Taylor error dT_q allows
struct SurrogateMixArgs {
    const double* A4;      // (36, 15) Mandel a4 map
    double sA, sB, sC;     // Sachs rheology parameters (rheo_params_tranisotropic, ef = +1)
    double snorm;          // isotropic Sachs normalisation (n' = 1), 1 for n' = -3
    double alpha;          // Taylor weight
    double tol;            // relative error bound of Eij
    double dT[6];          // calibrated relative error of the Taylor term per component (0: no bound)
};

// fused inference: state -> features -> network -> Eij (P, 6), Vi (P, 3)
struct SurrogateInferArgs {
    SurrogateFeatArgs fa;      // state, strain rate, envelope / audit flags (fa.feat unused)
    SurrogateMixArgs mix;      // Sachs term, Taylor weight, error bound
    const void* wf[4];         // bf16 weights in MFMA fragment order (MlpNet::wf)
    const float* bias[4];      // padded biases (256, 256, 256, 16)
    const float* norm;         // in_mean (36), in_std (36), out_mean (9), out_std (9)
    const float2* gtab;        // GELU table (launch_gelu_table)
    double* Eij;
    double* Vi;
    int* idx;                  // or null: flagged parcels appended here (arbitrary order)
    int* count;                // their number (zeroed by the caller)
};

hipError_t launch_features(const SurrogateFeatArgs& a, hipStream_t s);
hipError_t launch_surrogate_infer(const SurrogateInferArgs& a, hipStream_t s);
// the 1024-interval table of the normal CDF used by the fused kernel's GELU
hipError_t launch_gelu_table(float2* tab, hipStream_t s);
hipError_t launch_targets(int64_t P, const double* Eij, const double* Vi, float* T, hipStream_t s);
// merge of the work list (count on the device, at most cap items): fallback
// parcels take the plain values, audited ones record relative errors
hipError_t launch_merge(const int* count, int64_t cap, const int* idx, const int* flags, const double* Ep,
                        const double* Vp, double* Eij, double* Vi, int* status, double* errE, double* errV,
                        hipStream_t s);

// audit statistics of a merged work list, reduced on the device
struct AuditSums {
    long long n_fallback, n_audit, n_bound;
    double max_E, sum_E, max_V, sum_V;
};
hipError_t launch_audit_stats(const int* count, int64_t cap, const double* errE, const double* errV, AuditSums* out,
                              hipStream_t s);
// zeroes the work-list count (a kernel, so that the sequence can be captured
// in a graph)
hipError_t launch_reset_count(int* count, hipStream_t s);

}  // namespace sfg
