// Surrogate homogenization: feature extraction, training targets, the fused
// inference kernel, envelope / audit selection and the merge of the plain
// fallback.
//
// Features (36, fp32): Re and Im of the l <= 4 coefficients n_l^m / n_0^0 in
// specfab order (30) and the unit strain rate D / |D| in Mandel order (6).
// Outputs (9): log E_ij in the Cartesian frame (E11, E22, E33, E23, E13, E12;
// the eigenframe values of the plain path jump where a2 eigenvalues cross,
// the Cartesian ones are smooth in the state) and log of the elastic phase
// velocities along z sorted ascending (slower S, faster S, P).
#include <hip/hip_runtime.h>

#include "sfg_device.h"
#include "sfg_surrogate.h"

namespace sfg {

namespace {
unsigned nb(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
}

// symmetric 3x3 eigenvalues (cyclic Jacobi), sorted descending
__device__ inline void d_eigvals3(const double* M, double* lam) {
    double a[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a[i][j] = 0.5 * (M[3 * i + j] + M[3 * j + i]);
    for (int sweep = 0; sweep < 40; sweep++) {
        double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        if (off == 0.0) break;
#pragma unroll
        for (int pq = 0; pq < 3; pq++) {
            const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
            if (a[p][q] == 0.0) continue;
            double th = (a[q][q] - a[p][p]) / (2 * a[p][q]);
            double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
            double c = 1 / sqrt(t * t + 1), s = t * c;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double akp = a[k][p], akq = a[k][q];
                a[k][p] = c * akp - s * akq;
                a[k][q] = s * akp + c * akq;
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double apk = a[p][k], aqk = a[q][k];
                a[p][k] = c * apk - s * aqk;
                a[q][k] = s * apk + c * aqk;
            }
        }
    }
    double d0 = a[0][0], d1 = a[1][1], d2 = a[2][2], t;
    if (d1 > d0) { t = d0; d0 = d1; d1 = t; }
    if (d2 > d0) { t = d0; d0 = d2; d2 = t; }
    if (d2 > d1) { t = d1; d1 = d2; d2 = t; }
    lam[0] = d0; lam[1] = d1; lam[2] = d2;
}

// splitmix-style hash -> [0, 1)
__device__ inline double unit_hash(uint64_t v) {
    v += 0x9E3779B97F4A7C15ull;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    v ^= v >> 31;
    return (double)(v >> 11) * (1.0 / 9007199254740992.0);
}

// features of parcel p into f[36]; returns its SUR_* flags (0 without a.flags)
__device__ inline int parcel_features(const SurrogateFeatArgs& a, int64_t p, float* f) {
    double xq[15];
#pragma unroll
    for (int q = 0; q < 15; q++) xq[q] = a.x[(int64_t)q * a.P + p];
    const double x0 = xq[0];
    // complex l <= 4 coefficients from the real basis [Re n_l0, Re n_l1, Im n_l1, ...]
    int j = 0;
    for (int l = 0; l <= 4; l += 2) {
        const int base = (l == 0) ? 0 : (l == 2 ? 1 : 6);
        for (int m = -l; m <= l; m++, j++) {
            const int am = m < 0 ? -m : m;
            double re = xq[base + (am == 0 ? 0 : 2 * am - 1)];
            double im = am == 0 ? 0.0 : xq[base + 2 * am];
            if (m < 0) {
                const double s = (am & 1) ? -1.0 : 1.0;
                re *= s;
                im *= -s;
            }
            f[2 * j] = (float)(re / x0);
            f[2 * j + 1] = (float)(im / x0);
        }
    }
    // unit strain rate (Mandel)
    const double* G = a.ugrad + p * 9;
    double D[6] = {G[0], G[4], G[8], 0.5 * (G[5] + G[7]), 0.5 * (G[2] + G[6]), 0.5 * (G[1] + G[3])};
    double nn = D[0] * D[0] + D[1] * D[1] + D[2] * D[2] + 2 * (D[3] * D[3] + D[4] * D[4] + D[5] * D[5]);
    const double inv = nn > 0 ? 1.0 / sqrt(nn) : 0.0;
    const double r2 = sqrt(2.0);
    f[30] = (float)(D[0] * inv); f[31] = (float)(D[1] * inv); f[32] = (float)(D[2] * inv);
    f[33] = (float)(r2 * D[3] * inv); f[34] = (float)(r2 * D[4] * inv); f[35] = (float)(r2 * D[5] * inv);
    // envelope: sorted a2 eigenvalues inside the training box; audit sample
    int fl = 0;
    if (a.flags) {
        double a2v[6];
#pragma unroll
        for (int i = 0; i < 6; i++) {
            double s = 0;
#pragma unroll
            for (int q = 0; q < 6; q++) s += a.A2[i * 15 + q] * xq[q];
            a2v[i] = s / x0 + (i < 3 ? 1.0 / 3.0 : 0.0);
        }
        double A[9];
        A[0] = a2v[0]; A[4] = a2v[1]; A[8] = a2v[2];
        A[5] = A[7] = a2v[3] / r2; A[2] = A[6] = a2v[4] / r2; A[1] = A[3] = a2v[5] / r2;
        double lam[3];
        d_eigvals3(A, lam);
        for (int i = 0; i < 3; i++)
            if (!(lam[i] >= a.env_lo[i] && lam[i] <= a.env_hi[i])) fl |= SUR_FALLBACK;
        if (!(fl & SUR_FALLBACK) && unit_hash(a.seed ^ (uint64_t)p) < a.audit_frac) fl |= SUR_AUDIT;
    }
    return fl;
}

__global__ void features_kernel(SurrogateFeatArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.P) return;
    const int fl = parcel_features(a, p, a.feat + p * 36);
    if (a.flags) a.flags[p] = fl;
}

// training targets from plain results: log Eij (xyz) and log sorted velocities
__global__ void targets_kernel(int64_t P, const double* Eij, const double* Vi, float* T) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    for (int i = 0; i < 6; i++) T[p * 9 + i] = (float)log(Eij[p * 6 + i]);
    double v0 = Vi[p * 3], v1 = Vi[p * 3 + 1], v2 = Vi[p * 3 + 2], t;
    if (v1 < v0) { t = v0; v0 = v1; v1 = t; }
    if (v2 < v1) { t = v1; v1 = v2; v2 = t; }
    if (v1 < v0) { t = v0; v0 = v1; v1 = t; }
    T[p * 9 + 6] = (float)log(v0);
    T[p * 9 + 7] = (float)log(v1);
    T[p * 9 + 8] = (float)log(v2);
}

// work list of flagged parcels (fallback or audit); order is arbitrary
__global__ void select_kernel(int64_t P, const int* flags, int* idx, int* count) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    if (flags[p]) {
        int k = atomicAdd(count, 1);
        idx[k] = (int)p;
    }
}

// merge the plain results of the work list: fallback parcels take them,
// audited parcels keep the surrogate values and record relative errors
__global__ void merge_kernel(int64_t n, const int* idx, const int* flags, const double* Ep, const double* Vp,
                             double* Eij, double* Vi, int* status, double* errE, double* errV) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t p = idx[t];
    double v0 = Vp[t * 3], v1 = Vp[t * 3 + 1], v2 = Vp[t * 3 + 2], tt;
    if (v1 < v0) { tt = v0; v0 = v1; v1 = tt; }
    if (v2 < v1) { tt = v1; v1 = v2; v2 = tt; }
    if (v1 < v0) { tt = v0; v0 = v1; v1 = tt; }
    const double vs[3] = {v0, v1, v2};
    if (flags[p] & SUR_FALLBACK) {
        for (int i = 0; i < 6; i++) Eij[p * 6 + i] = Ep[t * 6 + i];
        for (int i = 0; i < 3; i++) Vi[p * 3 + i] = vs[i];
        errE[t] = -1.0;
        errV[t] = -1.0;
        if (status) atomicOr(status + p, 16);
    } else {
        double eE = 0, eV = 0;
        for (int i = 0; i < 6; i++) eE = fmax(eE, fabs(Eij[p * 6 + i] - Ep[t * 6 + i]) / fabs(Ep[t * 6 + i]));
        for (int i = 0; i < 3; i++) eV = fmax(eV, fabs(Vi[p * 3 + i] - vs[i]) / fabs(vs[i]));
        errE[t] = eE;
        errV[t] = eV;
        if (status) atomicOr(status + p, 32);
    }
}

// ------------------------------------------------------------ fused inference
//
// One kernel from the state to Eij / Vi: features -> 3 x (bf16 MFMA layer,
// bias, GELU) -> output layer -> de-normalisation and exp.  Each wave owns 64
// parcels and computes the transposed products Z^T = W^T H^T, i.e. hidden
// units along the MFMA M dimension and parcels along N (4 tiles of 16).  The
// accumulator of v_mfma_f32_16x16x32_bf16 (lane l holds C[4 (l/16) + i][l%16])
// then already holds, per lane, the values the next layer's B operand needs
// (B[8 (l/16) + j][l%16], j = 0..7) once the K index of the next weight matrix
// is permuted: physical k = 32 s + 8 g + j <-> hidden unit
// 32 s + 16 (j/4) + 4 g + j%4.  The weights are stored in that fragment order
// (pack_frag_kernel), so activations never leave registers and each 16-byte
// weight load feeds 4 MFMAs.

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// GELU(z) = z Phi(z), Phi(z) = 1/2 + zc Q(u) with zc = clamp(z, -4.25, 4.25),
// u = 2 zc^2 / 4.25^2 - 1 and Q a degree-8 polynomial fitted to the normal CDF
// (weighted least squares on Chebyshev nodes); |error| <= 3.7e-5 for |z| <= 15
// in FP32 (about 1 % of a bf16 step at 1), against erf-GELU.
__device__ inline float gelu_poly(float z) {
    const float zc = __builtin_amdgcn_fmed3f(z, -4.25f, 4.25f);
    const float u = fmaf(zc * zc, 0x1.c5894ep-4f, -1.0f);
    float q = 0x1.42cf3ap-9f;
    q = fmaf(q, u, -0x1.a3cc1ep-8f);
    q = fmaf(q, u, 0x1.193d8ap-7f);
    q = fmaf(q, u, -0x1.ce5844p-7f);
    q = fmaf(q, u, 0x1.a34ed4p-6f);
    q = fmaf(q, u, -0x1.425d56p-5f);
    q = fmaf(q, u, 0x1.c84d4ep-5f);
    q = fmaf(q, u, -0x1.4acc94p-4f);
    q = fmaf(q, u, 0x1.53d5eep-3f);
    return z * fmaf(zc, q, 0.5f);
}

// bias + GELU of a 256-row layer output, repacked as the next layer's B
// fragments (k step s, parcel tile t)
__device__ inline void gelu_repack(const f32x4 (&acc)[16][4], const float* __restrict__ bias, int g,
                                   bf16x8 (&B)[8][4]) {
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const f32x4 b0 = *(const f32x4*)(bias + 32 * s + 4 * g);
        const f32x4 b1 = *(const f32x4*)(bias + 32 * s + 16 + 4 * g);
#pragma unroll
        for (int t = 0; t < 4; t++) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                B[s][t][i] = (__bf16)gelu_poly(acc[2 * s][t][i] + b0[i]);
                B[s][t][4 + i] = (__bf16)gelu_poly(acc[2 * s + 1][t][i] + b1[i]);
            }
        }
    }
}

// 256 x 256 hidden layer: acc = W^T H^T with W in fragment order
__device__ inline void hidden_layer(const bf16x8* __restrict__ wf, const bf16x8 (&B)[8][4], int lane,
                                    f32x4 (&acc)[16][4]) {
#pragma unroll
    for (int m = 0; m < 16; m++)
#pragma unroll
        for (int t = 0; t < 4; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 8; s++)
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const bf16x8 A = wf[(m * 8 + s) * 64 + lane];
#pragma unroll
            for (int t = 0; t < 4; t++) acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[s][t], acc[m][t], 0, 0, 0);
        }
}

constexpr int FEAT_LD = 72;   // LDS row of one parcel's 64 padded features (+8: bank spread)

}  // namespace

__global__ void __launch_bounds__(256, 1) surrogate_infer_kernel(SurrogateInferArgs a) {
    __shared__ __bf16 fs[4][64][FEAT_LD];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t base = ((int64_t)blockIdx.x * 4 + wave) * 64;
    const int64_t P = a.fa.P;
    {   // features, one lane per parcel; lanes past P feed zeros
        const int64_t p = base + lane;
        float f[36];
        if (p < P) {
            const int fl = parcel_features(a.fa, p, f);
            if (a.fa.flags) a.fa.flags[p] = fl;
#pragma unroll
            for (int k = 0; k < 36; k++) fs[wave][lane][k] = (__bf16)((f[k] - a.norm[k]) / a.norm[36 + k]);
        } else {
#pragma unroll
            for (int k = 0; k < 36; k++) fs[wave][lane][k] = (__bf16)0.0f;
        }
#pragma unroll
        for (int k = 36; k < 64; k++) fs[wave][lane][k] = (__bf16)0.0f;
    }
    __syncthreads();
    const int g = lane >> 4, c = lane & 15;
    f32x4 acc[16][4];
    bf16x8 B[8][4];
    {   // input layer, K = 64 (36 features + zero padding), natural K order
        const bf16x8* wf = (const bf16x8*)a.wf[0];
#pragma unroll
        for (int m = 0; m < 16; m++)
#pragma unroll
            for (int t = 0; t < 4; t++) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 2; s++) {
            bf16x8 b[4];
#pragma unroll
            for (int t = 0; t < 4; t++) b[t] = *(const bf16x8*)&fs[wave][16 * t + c][32 * s + 8 * g];
#pragma unroll
            for (int m = 0; m < 16; m++) {
                const bf16x8 A = wf[(m * 2 + s) * 64 + lane];
#pragma unroll
                for (int t = 0; t < 4; t++) acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, b[t], acc[m][t], 0, 0, 0);
            }
        }
    }
    gelu_repack(acc, a.bias[0], g, B);
    hidden_layer((const bf16x8*)a.wf[1], B, lane, acc);
    gelu_repack(acc, a.bias[1], g, B);
    hidden_layer((const bf16x8*)a.wf[2], B, lane, acc);
    gelu_repack(acc, a.bias[2], g, B);
    // output layer: 16 rows (9 outputs + zero padding)
    f32x4 o[4];
#pragma unroll
    for (int t = 0; t < 4; t++) o[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bf16x8* wf3 = (const bf16x8*)a.wf[3];
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const bf16x8 A = wf3[s * 64 + lane];
#pragma unroll
        for (int t = 0; t < 4; t++) o[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[s][t], o[t], 0, 0, 0);
    }
    // outputs 4 g + i of parcel base + 16 t + c: de-normalise, exp
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int64_t q = base + 16 * t + c;
        if (q >= P) continue;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int oi = 4 * g + i;
            if (oi >= 9) continue;
            const float yn = o[t][i] + a.bias[3][oi];
            const float y = yn * a.norm[81 + oi] + a.norm[72 + oi];
            const double v = exp((double)y);
            if (oi < 6) a.Eij[q * 6 + oi] = v;
            else a.Vi[q * 3 + oi - 6] = v;
        }
    }
}

hipError_t launch_surrogate_infer(const SurrogateInferArgs& a, hipStream_t s) {
    if (a.fa.P == 0) return hipSuccess;
    hipLaunchKernelGGL(surrogate_infer_kernel, dim3(nb(a.fa.P, 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_features(const SurrogateFeatArgs& a, hipStream_t s) {
    if (a.P == 0) return hipSuccess;
    hipLaunchKernelGGL(features_kernel, dim3(nb(a.P, 128)), dim3(128), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_targets(int64_t P, const double* Eij, const double* Vi, float* T, hipStream_t s) {
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(targets_kernel, dim3(nb(P, 256)), dim3(256), 0, s, P, Eij, Vi, T);
    return hipGetLastError();
}

hipError_t launch_select(int64_t P, const int* flags, int* idx, int* count, hipStream_t s) {
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(select_kernel, dim3(nb(P, 256)), dim3(256), 0, s, P, flags, idx, count);
    return hipGetLastError();
}

hipError_t launch_merge(int64_t n, const int* idx, const int* flags, const double* Ep, const double* Vp, double* Eij,
                        double* Vi, int* status, double* errE, double* errV, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(merge_kernel, dim3(nb(n, 256)), dim3(256), 0, s, n, idx, flags, Ep, Vp, Eij, Vi, status,
                       errE, errV);
    return hipGetLastError();
}

}  // namespace sfg
