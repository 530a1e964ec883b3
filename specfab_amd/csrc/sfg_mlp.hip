// bf16 MFMA GEMM and the elementwise kernels of the surrogate network.
//
// GEMM: one 256-thread workgroup per 64 x 64 output tile; the four wave64s
// form a 2 x 2 grid and each owns a 32 x 32 block = 2 x 2 tiles of the
// gfx950 instruction v_mfma_f32_16x16x32_bf16 (FP32 accumulate).  Per K step
// of 32 the workgroup stages a 64 x 32 A tile and a 32 x 64 B tile in LDS
// (operands are read with general strides, so transposed operands of the
// backward pass need no copies).
//
// MFMA operand layout (16x16x32): lane l holds A[l % 16][8 (l / 16) + j] and
// B[8 (l / 16) + j][l % 16], j = 0..7; the accumulator lane l holds
// C[4 (l / 16) + i][l % 16], i = 0..3.  tests/test_gpu_mlp.py checks this
// layout against a CPU product.
#include <hip/hip_runtime.h>

#include "sfg_mlp.h"
#include "sfg_mlp_net.h"

namespace sfg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {
constexpr int TM = 64, TN = 64, TK = 32;
}

__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs a) {
    __shared__ __bf16 As[TM][TK + 8];   // +8: spread rows over LDS banks
    __shared__ __bf16 Bs[TN][TK + 8];   // stored transposed: Bs[n][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const __bf16 zero = (__bf16)0.0f;
    for (int k0 = 0; k0 < a.K; k0 += TK) {
        // stage: 64 x 32 elements each of A and B, 8 per thread
        for (int e = tid; e < TM * TK; e += 256) {
            int r = e / TK, c = e % TK;
            int gm = m0 + r, gk = k0 + c;
            As[r][c] = (gm < a.M && gk < a.K) ? a.A[(int64_t)gm * a.sam + (int64_t)gk * a.sak] : zero;
        }
        for (int e = tid; e < TN * TK; e += 256) {
            int nn = e % TN, c = e / TN;
            int gn = n0 + nn, gk = k0 + c;
            Bs[nn][c] = (gn < a.N && gk < a.K) ? a.B[(int64_t)gk * a.sbk + (int64_t)gn * a.sbn] : zero;
        }
        __syncthreads();
        const int r16 = lane & 15, kq = (lane >> 4) * 8;
        bf16x8 af[2], bfr[2];
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 8; j++) af[i][j] = As[wm + 16 * i + r16][kq + j];
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 8; j++) bfr[i][j] = Bs[wn + 16 * i + r16][kq + j];
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    const int c16 = lane & 15, rq = (lane >> 4) * 4;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int q = 0; q < 4; q++) {
                int gm = m0 + wm + 16 * i + rq + q, gn = n0 + wn + 16 * j + c16;
                if (gm < a.M && gn < a.N) {
                    float v = acc[i][j][q];
                    if (a.bias) v += a.bias[gn];
                    float* dst = a.C + (int64_t)gm * a.ldc + gn;
                    *dst = a.accumulate ? *dst + v : v;
                }
            }
}

// GELU (erf form) and its derivative
__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float dgelu(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__global__ void gelu_fwd_kernel(int64_t n, const float* z, __bf16* act) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) act[i] = (__bf16)gelu(z[i]);
}

__global__ void gelu_bwd_kernel(int64_t n, const float* dact, const float* z, float* dz, __bf16* dz_bf) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float d = dact[i] * dgelu(z[i]);
    dz[i] = d;
    dz_bf[i] = (__bf16)d;
}

__global__ void to_bf16_kernel(int64_t rows, int cols, int ld_in, int ld_out, const float* in, __bf16* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * ld_out) return;
    int64_t r = i / ld_out;
    int c = (int)(i % ld_out);
    out[i] = (__bf16)(c < cols ? in[r * ld_in + c] : 0.0f);
}

// column sums of a rows x cols fp32 matrix (bias gradients); one thread per column
__global__ void colsum_kernel(int rows, int cols, int ld, const float* in, float* out) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int r = 0; r < rows; r++) s += in[(int64_t)r * ld + c];
    out[c] = s;
}

__global__ void adam_kernel(int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2,
                            float eps, float bc1, float bc2) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float mi = b1 * m[i] + (1 - b1) * gi;
    float vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// dy = scale * (y - t); loss_acc[row] = sum_c (y - t)^2 (per-row partials)
__global__ void mse_grad_kernel(int rows, int cols, int ld, const float* y, const float* t, float scale, float* dy,
                                float* loss_acc) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    float s = 0.f;
    for (int c = 0; c < ld; c++) {
        float d = c < cols ? y[(int64_t)r * ld + c] - t[(int64_t)r * cols + c] : 0.f;
        dy[(int64_t)r * ld + c] = scale * d;
        s += d * d;
    }
    loss_acc[r] = s;
}

// weights of one layer (FP32 master, [in][out]) in the fragment order of the
// fused inference kernel: element ((m * KS + s) * 64 + lane) * 8 + j holds
// W[k][16 m + lane % 16] with k = 32 s + 8 (lane / 16) + j (input layer) or,
// for layers fed by a previous layer's accumulators, the permuted
// k = 32 s + 16 (j / 4) + 4 (lane / 16) + j % 4; zero outside in x out.
__global__ void pack_frag_kernel(int in, int out, int KS, int MT, int perm, const float* W, __bf16* wf) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)MT * KS * 512) return;
    const int j = (int)(e & 7), lane = (int)((e >> 3) & 63);
    const int64_t f = e >> 9;
    const int s = (int)(f % KS), m = (int)(f / KS);
    const int g = lane >> 4;
    const int k = perm ? 32 * s + 16 * (j >> 2) + 4 * g + (j & 3) : 32 * s + 8 * g + j;
    const int o = 16 * m + (lane & 15);
    wf[e] = (__bf16)((k < in && o < out) ? W[(int64_t)k * out + o] : 0.0f);
}

namespace {
unsigned nb(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
}

hipError_t launch_pack_frag(int in, int out, int in_pad, int out_pad, int perm, const float* W, __bf16* wf,
                            hipStream_t s) {
    const int KS = in_pad / 32, MT = out_pad / 16;
    const int64_t n = (int64_t)MT * KS * 512;
    hipLaunchKernelGGL(pack_frag_kernel, dim3(nb(n, 256)), dim3(256), 0, s, in, out, KS, MT, perm, W, wf);
    return hipGetLastError();
}

hipError_t launch_gemm_bf16(const GemmArgs& a, hipStream_t s) {
    if (a.M <= 0 || a.N <= 0) return hipSuccess;
    hipLaunchKernelGGL(gemm_bf16_kernel, dim3(nb(a.N, TN), nb(a.M, TM)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_gelu_fwd(int64_t n, const float* z, __bf16* act, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gelu_fwd_kernel, dim3(nb(n, 256)), dim3(256), 0, s, n, z, act);
    return hipGetLastError();
}

hipError_t launch_gelu_bwd(int64_t n, const float* dact, const float* z, float* dz, __bf16* dz_bf, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gelu_bwd_kernel, dim3(nb(n, 256)), dim3(256), 0, s, n, dact, z, dz, dz_bf);
    return hipGetLastError();
}

hipError_t launch_to_bf16(int64_t rows, int cols, int ld_in, int ld_out, const float* in, __bf16* out,
                          hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(to_bf16_kernel, dim3(nb(rows * ld_out, 256)), dim3(256), 0, s, rows, cols, ld_in, ld_out, in,
                       out);
    return hipGetLastError();
}

hipError_t launch_colsum(int rows, int cols, int ld, const float* in, float* out, hipStream_t s) {
    if (cols <= 0) return hipSuccess;
    hipLaunchKernelGGL(colsum_kernel, dim3(nb(cols, 64)), dim3(64), 0, s, rows, cols, ld, in, out);
    return hipGetLastError();
}

hipError_t launch_adam(int64_t n, float* w, const float* g, float* m, float* v, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(adam_kernel, dim3(nb(n, 256)), dim3(256), 0, s, n, w, g, m, v, lr, b1, b2, eps, bc1, bc2);
    return hipGetLastError();
}

hipError_t launch_mse_grad(int rows, int cols, int ld, const float* y, const float* t, float scale, float* dy,
                           float* loss_acc, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(mse_grad_kernel, dim3(nb(rows, 256)), dim3(256), 0, s, rows, cols, ld, y, t, scale, dy,
                       loss_acc);
    return hipGetLastError();
}

}  // namespace sfg

namespace sfg {

__global__ void normalize_in_kernel(int64_t B, const float* X, const float* norm, float* xin) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * MLP_IN_PAD) return;
    int64_t r = i / MLP_IN_PAD;
    int c = (int)(i % MLP_IN_PAD);
    xin[i] = c < MLP_IN ? (X[r * MLP_IN + c] - norm[c]) / norm[MLP_IN + c] : 0.0f;
}

__global__ void normalize_out_kernel(int64_t B, const float* T, const float* norm, float* tn) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * MLP_OUT) return;
    int c = (int)(i % MLP_OUT);
    tn[i] = (T[i] - norm[2 * MLP_IN + c]) / norm[2 * MLP_IN + MLP_OUT + c];
}

// yn (B x 16 padded) -> y (B x 9) raw units
__global__ void denormalize_out_kernel(int64_t B, const float* yn, const float* norm, float* y) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * MLP_OUT) return;
    int64_t r = i / MLP_OUT;
    int c = (int)(i % MLP_OUT);
    y[i] = yn[r * MLP_OUT_PAD + c] * norm[2 * MLP_IN + MLP_OUT + c] + norm[2 * MLP_IN + c];
}

hipError_t launch_normalize_in(int64_t B, const float* X, const float* norm, float* xin, hipStream_t s) {
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(normalize_in_kernel, dim3(nb(B * MLP_IN_PAD, 256)), dim3(256), 0, s, B, X, norm, xin);
    return hipGetLastError();
}

hipError_t launch_normalize_out(int64_t B, const float* T, const float* norm, float* tn, hipStream_t s) {
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(normalize_out_kernel, dim3(nb(B * MLP_OUT, 256)), dim3(256), 0, s, B, T, norm, tn);
    return hipGetLastError();
}

hipError_t launch_denormalize_out(int64_t B, const float* yn, const float* norm, float* y, hipStream_t s) {
    if (B <= 0) return hipSuccess;
    hipLaunchKernelGGL(denormalize_out_kernel, dim3(nb(B * MLP_OUT, 256)), dim3(256), 0, s, B, yn, norm, y);
    return hipGetLastError();
}

}  // namespace sfg
