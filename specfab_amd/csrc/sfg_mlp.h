// Learned homogenization: MLP 36 -> 256 -> 256 -> 256 -> 9 (GELU), bf16
// weights and activations, FP32 accumulation on the gfx950 matrix cores.
#pragma once
#ifdef SFG_HOST_EMUL
#include "sfg_emul.h"
#else
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

namespace sfg {

constexpr int MLP_IN = 36, MLP_H = 256, MLP_OUT = 9;
constexpr int MLP_IN_PAD = 64, MLP_OUT_PAD = 16;   // K and N padded to the MFMA tile
constexpr int MLP_NLAYER = 4;
// flattened FP32 parameters: W1 (36x256, [in][out]), b1, W2, b2, W3, b3, W4 (256x9), b4
constexpr int64_t MLP_NPARAM = (int64_t)MLP_IN * MLP_H + MLP_H + 2 * ((int64_t)MLP_H * MLP_H + MLP_H) +
                               (int64_t)MLP_H * MLP_OUT + MLP_OUT;

// C (M x N, fp32, row stride ldc) = A (M x K) * B (K x N) [+ bias (N)], bf16
// operands addressed with general strides (so transposed operands cost
// nothing): A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn].
struct GemmArgs {
    int M, N, K;
    const __bf16* A;
    int64_t sam, sak;
    const __bf16* B;
    int64_t sbk, sbn;
    float* C;
    int64_t ldc;
    const float* bias;   // or null
    int accumulate;      // 1: C += A*B
};

hipError_t launch_gemm_bf16(const GemmArgs& a, hipStream_t s);

// weights (FP32 [in][out]) -> bf16 MFMA fragment order of the fused inference
// kernel (in_pad a multiple of 32, out_pad of 16); perm: permuted K order of
// layers fed by accumulators
hipError_t launch_pack_frag(int in, int out, int in_pad, int out_pad, int perm, const float* W, __bf16* wf,
                            hipStream_t s);

// elementwise helpers
hipError_t launch_gelu_fwd(int64_t n, const float* z, __bf16* act, hipStream_t s);
hipError_t launch_gelu_bwd(int64_t n, const float* dact, const float* z, float* dz, __bf16* dz_bf, hipStream_t s);
hipError_t launch_to_bf16(int64_t rows, int cols, int ld_in, int ld_out, const float* in, __bf16* out,
                          hipStream_t s);
hipError_t launch_colsum(int rows, int cols, int ld, const float* in, float* out, hipStream_t s);
hipError_t launch_adam(int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, hipStream_t s);
hipError_t launch_mse_grad(int rows, int cols, int ld, const float* y, const float* t, float scale, float* dy,
                           float* loss_acc, hipStream_t s);

}  // namespace sfg
