// Host side of the surrogate network: parameter storage (FP32 master copy
// plus bf16 working copy), training step (forward, backward, Adam) and batched
// inference, all as launches of the bf16 MFMA GEMM and elementwise kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/specfab_gpu.h"
#include "sfg_mlp.h"
#include "sfg_mlp_net.h"
#include "sfg_surrogate.h"

namespace sfg {

namespace {

struct Layer { int in, out, in_pad, out_pad; int64_t w_off, b_off; };

Layer layer(int i) {
    static const int dims[5] = {MLP_IN, MLP_H, MLP_H, MLP_H, MLP_OUT};
    Layer L{};
    L.in = dims[i];
    L.out = dims[i + 1];
    L.in_pad = (i == 0) ? MLP_IN_PAD : MLP_H;
    L.out_pad = (i == 3) ? MLP_OUT_PAD : MLP_H;
    int64_t off = 0;
    for (int j = 0; j < i; j++) off += (int64_t)dims[j] * dims[j + 1] + dims[j + 1];
    L.w_off = off;
    L.b_off = off + (int64_t)L.in * L.out;
    return L;
}

}  // namespace

MlpNet::~MlpNet() {
    for (void* p : owned) (void)hipFree(p);
}

hipError_t MlpNet::alloc(void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) owned.push_back(*p);
    return e;
}

hipError_t MlpNet::init(int device_, uint64_t seed) {
    device = device_;
    hipError_t e;
    if ((e = hipSetDevice(device)) != hipSuccess) return e;
    const size_t np = (size_t)MLP_NPARAM;
    if ((e = alloc((void**)&w, np * 4)) != hipSuccess) return e;
    if ((e = alloc((void**)&g, np * 4)) != hipSuccess) return e;
    if ((e = alloc((void**)&m, np * 4)) != hipSuccess) return e;
    if ((e = alloc((void**)&v, np * 4)) != hipSuccess) return e;
    if ((e = alloc((void**)&norm, 2 * (MLP_IN + MLP_OUT) * 4)) != hipSuccess) return e;
    if ((e = alloc((void**)&gtab, 1024 * sizeof(float2))) != hipSuccess) return e;
    if ((e = launch_gelu_table(gtab, nullptr)) != hipSuccess) return e;
    for (int i = 0; i < MLP_NLAYER; i++) {
        Layer L = layer(i);
        if ((e = alloc((void**)&wb[i], (size_t)L.in_pad * L.out_pad * 2)) != hipSuccess) return e;
        if ((e = alloc((void**)&bias_pad[i], (size_t)L.out_pad * 4)) != hipSuccess) return e;
        if ((e = alloc((void**)&wf[i], (size_t)L.in_pad * L.out_pad * 2)) != hipSuccess) return e;
    }
    // He-uniform init of W, zero biases; identity normalisation
    std::mt19937_64 rng(seed);
    std::vector<float> h(np, 0.0f);
    for (int i = 0; i < MLP_NLAYER; i++) {
        Layer L = layer(i);
        float lim = std::sqrt(6.0f / (float)L.in);
        std::uniform_real_distribution<float> U(-lim, lim);
        for (int64_t k = 0; k < (int64_t)L.in * L.out; k++) h[L.w_off + k] = U(rng);
    }
    std::vector<float> nm(2 * (MLP_IN + MLP_OUT));
    for (int i = 0; i < MLP_IN; i++) { nm[i] = 0.f; nm[MLP_IN + i] = 1.f; }
    for (int i = 0; i < MLP_OUT; i++) { nm[2 * MLP_IN + i] = 0.f; nm[2 * MLP_IN + MLP_OUT + i] = 1.f; }
    norm_host = nm;
    if ((e = hipMemcpy(norm, nm.data(), nm.size() * 4, hipMemcpyHostToDevice)) != hipSuccess) return e;
    if ((e = hipMemset(m, 0, np * 4)) != hipSuccess) return e;
    if ((e = hipMemset(v, 0, np * 4)) != hipSuccess) return e;
    step = 0;
    return set_params(h.data());
}

hipError_t MlpNet::set_params(const float* host) {
    hipError_t e = hipMemcpy(w, host, (size_t)MLP_NPARAM * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    e = refresh_bf16(nullptr);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(nullptr);
}

// bf16 working copies (padded [in_pad][out_pad]) and padded biases from the
// FP32 master weights
hipError_t MlpNet::refresh_bf16(hipStream_t s) {
    for (int i = 0; i < MLP_NLAYER; i++) {
        Layer L = layer(i);
        hipError_t e = launch_to_bf16(L.in, L.out, L.out, L.out_pad, w + L.w_off, wb[i], s);
        if (e != hipSuccess) return e;
        // padding rows in..in_pad-1 (first layer) are zero
        if (L.in_pad > L.in) {
            e = hipMemsetAsync(wb[i] + (size_t)L.in * L.out_pad, 0, (size_t)(L.in_pad - L.in) * L.out_pad * 2, s);
            if (e != hipSuccess) return e;
        }
        e = hipMemsetAsync(bias_pad[i], 0, (size_t)L.out_pad * 4, s);
        if (e != hipSuccess) return e;
        e = hipMemcpyAsync(bias_pad[i], w + L.b_off, (size_t)L.out * 4, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
        e = launch_pack_frag(L.in, L.out, L.in_pad, L.out_pad, i > 0 ? 1 : 0, w + L.w_off, wf[i], s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t MlpNet::ensure_batch(int64_t B) {
    if (B <= cap) return hipSuccess;
    // queued work may still use the old buffers
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return e;
    for (void* p : batch_bufs) {
        for (size_t i = 0; i < owned.size(); i++)
            if (owned[i] == p) { owned.erase(owned.begin() + i); break; }
        (void)hipFree(p);
    }
    batch_bufs.clear();
    cap = 0;
    auto A = [&](void** p, size_t bytes) {
        hipError_t r = alloc(p, bytes);
        if (r == hipSuccess) batch_bufs.push_back(*p);
        return r;
    };
    const size_t Bz = (size_t)B;
    if ((e = A((void**)&xin, Bz * MLP_IN_PAD * 4)) != hipSuccess) return e;
    if ((e = A((void**)&xin_b, Bz * MLP_IN_PAD * 2)) != hipSuccess) return e;
    for (int i = 0; i < 3; i++) {
        if ((e = A((void**)&z[i], Bz * MLP_H * 4)) != hipSuccess) return e;
        if ((e = A((void**)&act[i], Bz * MLP_H * 2)) != hipSuccess) return e;
    }
    if ((e = A((void**)&yout, Bz * MLP_OUT_PAD * 4)) != hipSuccess) return e;
    if ((e = A((void**)&dy, Bz * MLP_OUT_PAD * 4)) != hipSuccess) return e;
    if ((e = A((void**)&dy_b, Bz * MLP_OUT_PAD * 2)) != hipSuccess) return e;
    if ((e = A((void**)&dact, Bz * MLP_H * 4)) != hipSuccess) return e;
    if ((e = A((void**)&dz, Bz * MLP_H * 4)) != hipSuccess) return e;
    if ((e = A((void**)&dz_b, Bz * MLP_H * 2)) != hipSuccess) return e;
    if ((e = A((void**)&tnorm, Bz * MLP_OUT * 4)) != hipSuccess) return e;
    if ((e = A((void**)&loss_rows, Bz * 4)) != hipSuccess) return e;
    cap = B;
    return hipSuccess;
}

// forward pass of B rows; X raw features (B x 36 fp32) -> normalised outputs
// in yout (B x 16, columns 9..15 zero)
hipError_t MlpNet::forward(int64_t B, const float* X, hipStream_t s) {
    hipError_t e = launch_normalize_in(B, X, norm, xin, s);
    if (e != hipSuccess) return e;
    if ((e = launch_to_bf16(B, MLP_IN_PAD, MLP_IN_PAD, MLP_IN_PAD, xin, xin_b, s)) != hipSuccess) return e;
    const __bf16* in = xin_b;
    int in_pad = MLP_IN_PAD;
    for (int i = 0; i < MLP_NLAYER; i++) {
        Layer L = layer(i);
        GemmArgs g{};
        g.M = (int)B; g.N = L.out_pad; g.K = in_pad;
        g.A = in; g.sam = in_pad; g.sak = 1;
        g.B = wb[i]; g.sbk = L.out_pad; g.sbn = 1;
        g.C = (i < 3) ? z[i] : yout; g.ldc = L.out_pad;
        g.bias = bias_pad[i];
        if ((e = launch_gemm_bf16(g, s)) != hipSuccess) return e;
        if (i < 3) {
            if ((e = launch_gelu_fwd(B * MLP_H, z[i], act[i], s)) != hipSuccess) return e;
            in = act[i];
            in_pad = MLP_H;
        }
    }
    return hipSuccess;
}

// loss and gradients (into g) of a batch; T raw targets (B x 9)
hipError_t MlpNet::backward(int64_t B, const float* T, hipStream_t s) {
    hipError_t e = launch_normalize_out(B, T, norm, tnorm, s);
    if (e != hipSuccess) return e;
    const float scale = 2.0f / (float)(B * MLP_OUT);
    if ((e = launch_mse_grad((int)B, MLP_OUT, MLP_OUT_PAD, yout, tnorm, scale, dy, loss_rows, s)) != hipSuccess)
        return e;
    if ((e = launch_to_bf16(B, MLP_OUT_PAD, MLP_OUT_PAD, MLP_OUT_PAD, dy, dy_b, s)) != hipSuccess) return e;
    const __bf16* dnext = dy_b;   // gradient w.r.t. the current layer's output (bf16)
    const float* dnext_f = dy;
    int dnext_ld = MLP_OUT_PAD;
    for (int i = MLP_NLAYER - 1; i >= 0; i--) {
        Layer L = layer(i);
        const __bf16* ain = (i == 0) ? xin_b : act[i - 1];
        const int ain_ld = (i == 0) ? MLP_IN_PAD : MLP_H;
        // dW (in x out) = A_in^T (in x B) * dOut (B x out)
        GemmArgs gw{};
        gw.M = L.in; gw.N = L.out; gw.K = (int)B;
        gw.A = ain; gw.sam = 1; gw.sak = ain_ld;
        gw.B = dnext; gw.sbk = dnext_ld; gw.sbn = 1;
        gw.C = g + L.w_off; gw.ldc = L.out;
        if ((e = launch_gemm_bf16(gw, s)) != hipSuccess) return e;
        // db = column sums of dOut (fp32)
        if ((e = launch_colsum((int)B, L.out, dnext_ld, dnext_f, g + L.b_off, s)) != hipSuccess) return e;
        if (i == 0) break;
        // dA_in (B x in) = dOut (B x out) * W^T (out x in)
        GemmArgs ga{};
        ga.M = (int)B; ga.N = L.in; ga.K = L.out_pad;
        ga.A = dnext; ga.sam = dnext_ld; ga.sak = 1;
        ga.B = wb[i]; ga.sbk = 1; ga.sbn = L.out_pad;
        ga.C = dact; ga.ldc = MLP_H;
        if ((e = launch_gemm_bf16(ga, s)) != hipSuccess) return e;
        if ((e = launch_gelu_bwd(B * MLP_H, dact, z[i - 1], dz, dz_b, s)) != hipSuccess) return e;
        dnext = dz_b;
        dnext_f = dz;
        dnext_ld = MLP_H;
    }
    return hipSuccess;
}

hipError_t MlpNet::adam(float lr, hipStream_t s) {
    step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.0f - std::pow(b1, (float)step), bc2 = 1.0f - std::pow(b2, (float)step);
    hipError_t e = launch_adam(MLP_NPARAM, w, g, m, v, lr, b1, b2, 1e-8f, bc1, bc2, s);
    if (e != hipSuccess) return e;
    return refresh_bf16(s);
}

}  // namespace sfg
