// Device-resident surrogate network (see sfg_mlp_host.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "sfg_mlp.h"

namespace sfg {

struct MlpNet {
    int device = 0;
    int64_t step = 0;
    float* w = nullptr;             // FP32 master parameters (MLP_NPARAM)
    float* g = nullptr;             // gradients
    float *m = nullptr, *v = nullptr;  // Adam moments
    float* norm = nullptr;          // in_mean(36), in_std(36), out_mean(9), out_std(9)
    std::vector<float> norm_host;
    __bf16* wb[MLP_NLAYER] = {};    // bf16 weights [in_pad][out_pad]
    float* bias_pad[MLP_NLAYER] = {};
    __bf16* wf[MLP_NLAYER] = {};    // bf16 weights in the fused inference kernel's fragment order
    float2* gtab = nullptr;         // GELU table of the fused inference kernel
    // batch buffers
    int64_t cap = 0;
    float* xin = nullptr;           // normalised input (B x 64)
    __bf16* xin_b = nullptr;
    float* z[3] = {};               // pre-activations (B x 256)
    __bf16* act[3] = {};            // GELU outputs (B x 256)
    float* yout = nullptr;          // normalised outputs (B x 16)
    float *dy = nullptr, *dact = nullptr, *dz = nullptr, *tnorm = nullptr, *loss_rows = nullptr;
    __bf16 *dy_b = nullptr, *dz_b = nullptr;
    std::vector<void*> owned, batch_bufs;
    // envelope of the training states: sorted a2 eigenvalues lo/hi
    double env_lo[3] = {0, 0, 0}, env_hi[3] = {1, 1, 1};
    // calibrated relative error of the six Taylor-term outputs (0: no bound)
    double taylor_err[6] = {0, 0, 0, 0, 0, 0};
    bool trained = false;

    ~MlpNet();
    hipError_t alloc(void** p, size_t bytes);
    hipError_t init(int device, uint64_t seed);
    hipError_t set_params(const float* host);
    hipError_t refresh_bf16(hipStream_t s);
    hipError_t ensure_batch(int64_t B);
    hipError_t forward(int64_t B, const float* X, hipStream_t s);
    hipError_t backward(int64_t B, const float* T, hipStream_t s);
    hipError_t adam(float lr, hipStream_t s);
};

hipError_t launch_normalize_in(int64_t B, const float* X, const float* norm, float* xin, hipStream_t s);
hipError_t launch_normalize_out(int64_t B, const float* T, const float* norm, float* tn, hipStream_t s);
hipError_t launch_denormalize_out(int64_t B, const float* yn, const float* norm, float* y, hipStream_t s);

}  // namespace sfg
