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

// eigenvalues of a symmetric 3x3 matrix, sorted descending: the
// trigonometric solution of the characteristic cubic (Smith, Commun. ACM 4,
// 1961), A = q I + p B with q = tr A / 3, p^2 = |A - q I|^2 / 6 and
// lambda = q + 2 p cos(acos(det B / 2) / 3 + 2 pi k / 3)
__device__ inline void d_eigvals3(const double* A, double* lam) {
    const double a00 = A[0], a11 = A[4], a22 = A[8];
    const double a01 = 0.5 * (A[1] + A[3]), a02 = 0.5 * (A[2] + A[6]), a12 = 0.5 * (A[5] + A[7]);
    const double q = (a00 + a11 + a22) / 3.0;
    const double b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
    const double p2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0 * (a01 * a01 + a02 * a02 + a12 * a12);
    const double p = sqrt(p2 / 6.0);
    if (!(p > 0.0)) {
        lam[0] = lam[1] = lam[2] = q;
        return;
    }
    const double ip = 1.0 / p;
    const double c00 = b00 * ip, c11 = b11 * ip, c22 = b22 * ip, c01 = a01 * ip, c02 = a02 * ip, c12 = a12 * ip;
    double r = 0.5 * (c00 * (c11 * c22 - c12 * c12) - c01 * (c01 * c22 - c12 * c02) + c02 * (c01 * c12 - c11 * c02));
    r = fmin(1.0, fmax(-1.0, r));
    const double phi = acos(r) / 3.0;
    const double e0 = q + 2.0 * p * cos(phi);
    const double e2 = q + 2.0 * p * cos(phi + 2.0943951023931957);
    lam[0] = e0;
    lam[1] = 3.0 * q - e0 - e2;
    lam[2] = e2;
}

// splitmix-style hash -> [0, 1)
__device__ inline double unit_hash(uint64_t v) {
    v += 0x9E3779B97F4A7C15ull;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    v ^= v >> 31;
    return (double)(v >> 11) * (1.0 / 9007199254740992.0);
}

// Sachs enhancement factors of the Cartesian axes (E11, E22, E33, E23, E13,
// E12): the alpha = 0 part of Eij_tranisotropic with (e1, e2, e3) = (x, y, z)
// (src/enhancementfactors.f90:23-69 with src/homogenizations.f90:69-116),
// written out for the axis stress states: tau = I/3 - e_i e_i (normal) and
// e_i e_j + e_j e_i (shear) reduce a4 : tau to the Mandel a4 entries a4v[i][j],
// i, j < 3, and a4v[q][q], q = 3..5, and tau.a2 + a2.tau to a2 diagonals
__device__ inline void sachs_xyz(const SurrogateMixArgs& m, const double* xq, double ix0, const double* a2v,
                                 double* S) {
    double a4r[3][3], a4d[3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double s = 0;
#pragma unroll
            for (int q = 0; q < 15; q++) s += m.A4[(6 * i + j) * 15 + q] * xq[q];
            a4r[i][j] = s * ix0;
        }
#pragma unroll
    for (int k = 3; k < 6; k++) {
        double s = 0;
#pragma unroll
        for (int q = 0; q < 15; q++) s += m.A4[(7 * k) * 15 + q] * xq[q];
        a4d[k - 3] = s * ix0;
    }
    const double tr3 = (a2v[0] + a2v[1] + a2v[2]) / 3.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        // tau_ii = tau : e_i e_i = -2/3; a2 : tau = tr a2 / 3 - a2_ii; (tau.a2 + a2.tau)_ii = -4/3 a2_ii
        double a4t = 0;
#pragma unroll
        for (int j = 0; j < 3; j++) a4t += a4r[i][j] * ((i == j ? -2.0 : 1.0) / 3.0);
        const double lin = -2.0 / 3.0 - m.sA * (tr3 - a2v[i]) + m.sB * a4t + m.sC * (-4.0 / 3.0 * a2v[i]);
        S[i] = lin / (m.snorm * (-2.0 / 3.0));
    }
    // shear (i, j) = (y, z), (x, z), (x, y): tau_ij = 1, (a4 : tau)_ij = a4v[q][q],
    // (tau.a2 + a2.tau)_ij = a2_ii + a2_jj, no isotropic (sA) part off the diagonal
    const int pi[3] = {1, 0, 0}, pj[3] = {2, 2, 1};
#pragma unroll
    for (int k = 0; k < 3; k++)
        S[3 + k] = (1.0 + m.sB * a4d[k] + m.sC * (a2v[pi[k]] + a2v[pj[k]])) / m.snorm;
}

// features of parcel p into f[36]; returns its SUR_* flags (0 without
// a.flags); with `mix`, also the Sachs term of its Cartesian Eij -> sachs[6]
__device__ inline int parcel_features(const SurrogateFeatArgs& a, int64_t p, float* f,
                                      const SurrogateMixArgs* mix = nullptr, float* sachs = nullptr) {
    double xq[15];
#pragma unroll
    for (int q = 0; q < 15; q++) xq[q] = a.x[(int64_t)q * a.P + p];
    const double x0 = xq[0], ix0 = 1.0 / x0;
    const double r2 = sqrt(2.0);
    // a2 first: the envelope test and the Sachs term are done (and the Sachs
    // term stored) before the 36 features are formed, so that they are not
    // live at the same time
    int fl = 0;
    if (a.flags || mix) {
        double a2v[6];
#pragma unroll
        for (int i = 0; i < 6; i++) {
            double s = 0;
#pragma unroll
            for (int q = 0; q < 6; q++) s += a.A2[i * 15 + q] * xq[q];
            a2v[i] = s * ix0 + (i < 3 ? 1.0 / 3.0 : 0.0);
        }
        if (mix) {
            double S[6];
            sachs_xyz(*mix, xq, ix0, a2v, S);
#pragma unroll
            for (int i = 0; i < 6; i++) sachs[i] = (float)S[i];
        }
        // envelope: sorted a2 eigenvalues inside the training box; audit sample
        if (a.flags) {
            double A[9];
            A[0] = a2v[0]; A[4] = a2v[1]; A[8] = a2v[2];
            A[5] = A[7] = a2v[3] / r2; A[2] = A[6] = a2v[4] / r2; A[1] = A[3] = a2v[5] / r2;
            double lam[3];
            d_eigvals3(A, lam);
            for (int i = 0; i < 3; i++)
                if (!(lam[i] >= a.env_lo[i] && lam[i] <= a.env_hi[i])) fl |= SUR_FALLBACK;
            if (!(fl & SUR_FALLBACK) && unit_hash(a.seed ^ (uint64_t)p) < a.audit_frac) fl |= SUR_AUDIT;
        }
    }
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
            f[2 * j] = (float)(re * ix0);
            f[2 * j + 1] = (float)(im * ix0);
        }
    }
    // unit strain rate (Mandel)
    const double* G = a.ugrad + p * 9;
    double D[6] = {G[0], G[4], G[8], 0.5 * (G[5] + G[7]), 0.5 * (G[2] + G[6]), 0.5 * (G[1] + G[3])};
    double nn = D[0] * D[0] + D[1] * D[1] + D[2] * D[2] + 2 * (D[3] * D[3] + D[4] * D[4] + D[5] * D[5]);
    const double inv = nn > 0 ? 1.0 / sqrt(nn) : 0.0;
    f[30] = (float)(D[0] * inv); f[31] = (float)(D[1] * inv); f[32] = (float)(D[2] * inv);
    f[33] = (float)(r2 * D[3] * inv); f[34] = (float)(r2 * D[4] * inv); f[35] = (float)(r2 * D[5] * inv);
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

// merge the plain results of the work list: fallback parcels take them,
// audited parcels keep the surrogate values and record relative errors
__device__ __forceinline__ void merge_item(int64_t t, const int* idx, const int* flags, const double* Ep,
                                           const double* Vp, double* Eij, double* Vi, int* status, double* errE,
                                           double* errV) {
    const int64_t p = idx[t];
    double v0 = Vp[t * 3], v1 = Vp[t * 3 + 1], v2 = Vp[t * 3 + 2], tt;
    if (v1 < v0) { tt = v0; v0 = v1; v1 = tt; }
    if (v2 < v1) { tt = v1; v1 = v2; v2 = tt; }
    if (v1 < v0) { tt = v0; v0 = v1; v1 = tt; }
    const double vs[3] = {v0, v1, v2};
    const int fl = flags[p];
    if (fl & (SUR_FALLBACK | SUR_BOUND)) {
        for (int i = 0; i < 6; i++) Eij[p * 6 + i] = Ep[t * 6 + i];
        for (int i = 0; i < 3; i++) Vi[p * 3 + i] = vs[i];
        // -2 marks the parcels that only the error bound sent to the plain path
        errE[t] = (fl & SUR_FALLBACK) ? -1.0 : -2.0;
        errV[t] = -1.0;
        if (status) atomicOr(status + p, ((fl & SUR_FALLBACK) ? 16 : 0) | ((fl & SUR_BOUND) ? 64 : 0));
    } else {
        double eE = 0, eV = 0;
        for (int i = 0; i < 6; i++) eE = fmax(eE, fabs(Eij[p * 6 + i] - Ep[t * 6 + i]) / fabs(Ep[t * 6 + i]));
        for (int i = 0; i < 3; i++) eV = fmax(eV, fabs(Vi[p * 3 + i] - vs[i]) / fabs(vs[i]));
        errE[t] = eE;
        errV[t] = eV;
        if (status) atomicOr(status + p, 32);
    }
}

__global__ void merge_kernel(const int* count, int64_t cap, const int* idx, const int* flags, const double* Ep,
                             const double* Vp, double* Eij, double* Vi, int* status, double* errE, double* errV) {
    const int64_t n = *count < cap ? *count : cap;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride)
        merge_item(t, idx, flags, Ep, Vp, Eij, Vi, status, errE, errV);
}

// one workgroup of 256 threads: counts, maxima and sums of the merged list
// (errE < 0 marks a fallback parcel)
__global__ void __launch_bounds__(256) audit_stats_kernel(const int* count, int64_t cap, const double* errE,
                                                          const double* errV, AuditSums* out) {
    const int64_t n = *count < cap ? *count : cap;
    long long nf = 0, na = 0, nb = 0;
    double mE = 0, sE = 0, mV = 0, sV = 0;
    for (int64_t t = threadIdx.x; t < n; t += 256) {
        const double e = errE[t], v = errV[t];
        if (e < 0) {
            nf++;
            if (e < -1.5) nb++;
        } else {
            na++;
            mE = fmax(mE, e);
            sE += e;
            mV = fmax(mV, v);
            sV += v;
        }
    }
    __shared__ long long snf[256], sna[256], snb[256];
    __shared__ double smE[256], ssE[256], smV[256], ssV[256];
    const int i = threadIdx.x;
    snf[i] = nf; sna[i] = na; snb[i] = nb; smE[i] = mE; ssE[i] = sE; smV[i] = mV; ssV[i] = sV;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (i < h) {
            snf[i] += snf[i + h];
            sna[i] += sna[i + h];
            snb[i] += snb[i + h];
            smE[i] = fmax(smE[i], smE[i + h]);
            ssE[i] += ssE[i + h];
            smV[i] = fmax(smV[i], smV[i + h]);
            ssV[i] += ssV[i + h];
        }
        __syncthreads();
    }
    if (i == 0) {
        out->n_fallback = snf[0];
        out->n_audit = sna[0];
        out->n_bound = snb[0];
        out->max_E = smE[0];
        out->sum_E = ssE[0];
        out->max_V = smV[0];
        out->sum_V = ssV[0];
    }
}

__global__ void reset_count_kernel(int* count) { *count = 0; }

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

// GELU(z) = z Phi(z) with Phi interpolated linearly in a table of 1024
// intervals on [-5, 5] (LDS, built by each workgroup from erfc): nodes
// z_k = (k - 512) / 102.4, entry k = (Phi(z_k), Phi(z_k+1) - Phi(z_k)).
// Interpolation error of Phi <= h^2 / 8 max|Phi''| = 3e-6 (of GELU <= 4e-6 on
// [-5, 5]); beyond +-5 Phi is held at its end values (1 - Phi(5) = 2.9e-7).
constexpr int GTAB_N = 1024;
constexpr float GTAB_SCALE = 102.4f, GTAB_OFF = 512.0f;

__device__ inline float gtab_phi(int k) {
    const float z = ((float)k - GTAB_OFF) / GTAB_SCALE;
    return 0.5f * erfcf(-z * 0.70710678118654752f);
}

__device__ inline float gelu_tab(const float2* __restrict__ tab, float z) {
    const float t = __builtin_amdgcn_fmed3f(fmaf(z, GTAB_SCALE, GTAB_OFF), 0.0f, 0x1.fffffep+9f);
    const float2 e = tab[(int)t];
    return z * fmaf(e.y, __builtin_amdgcn_fractf(t), e.x);
}

// GELU of a 256-row layer output (bias already in the accumulator), repacked
// as the next layer's B fragments (k step s, parcel tile t)
template <int NT>
__device__ inline void gelu_repack(const float2* __restrict__ tab, const f32x4 (&acc)[16][NT],
                                   bf16x8 (&B)[8][NT]) {
#pragma unroll
    for (int s = 0; s < 8; s++)
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                B[s][t][i] = (__bf16)gelu_tab(tab, acc[2 * s][t][i]);
                B[s][t][4 + i] = (__bf16)gelu_tab(tab, acc[2 * s + 1][t][i]);
            }
}

// bias as the initial accumulator: element i of tile (m, t) is row 16 m + 4 g + i
template <int NT>
__device__ inline void acc_bias(const float* __restrict__ bias, int g, f32x4 (&acc)[16][NT]) {
#pragma unroll
    for (int m = 0; m < 16; m++) {
        const f32x4 b = *(const f32x4*)(bias + 16 * m + 4 * g);
#pragma unroll
        for (int t = 0; t < NT; t++) acc[m][t] = b;
    }
}

constexpr int WSLICE = 16 * 64;   // weight fragments of one k step for the 16 M tiles

// workgroup barrier without the memory fence of __syncthreads() (whose
// vmcnt(0) would also wait for the weight slices still in flight): the
// caller's s_waitcnt completes the LDS traffic that must be visible, and the
// empty asm statements keep the compiler from moving memory accesses across
__device__ inline void wg_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// slice s of a layer's weights -> LDS by direct global-to-LDS loads (no
// registers): the wave's 64 lanes move 64 consecutive 16-byte fragments, lane
// i to LDS base + 16 i, four times
template <int KS>
__device__ inline void stage_slice(const bf16x8* __restrict__ wf, int s, bf16x8* buf, int tid) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int e = tid + 256 * q;
        __builtin_amdgcn_global_load_lds((const void*)(wf + ((e >> 6) * KS + s) * 64 + (e & 63)),
                                         (__attribute__((address_space(3))) void*)(buf + (e & ~63)), 16, 0, 0);
    }
}

// acc += W^T H^T for a layer of 256 rows and KS k steps (fragment (m, s, lane)
// at wf[(m KS + s) 64 + lane]); B fragments from bfrag(s, t).  The weights of
// a k step are staged once per workgroup in LDS and shared by the four waves;
// triple buffered: slices s + 1 and s + 2 are in flight during the MFMAs of
// step s.
// G0: number of k steps of the earlier layers; step s of this layer uses
// buffer (G0 + s) % 3, so that across layers too a buffer is refilled only
// after the barrier that follows every wave's last read of it.
template <int KS, int G0, int NT, class BF>
__device__ inline void mfma_layer(const bf16x8* __restrict__ wf, BF&& bfrag, f32x4 (&acc)[16][NT],
                                  bf16x8* const (&wst)[3], int tid, int lane) {
    stage_slice<KS>(wf, 0, wst[G0 % 3], tid);
    stage_slice<KS>(wf, 1, wst[(G0 + 1) % 3], tid);
#pragma unroll
    for (int s = 0; s < KS; s++) {
        const bf16x8* buf = wst[(G0 + s) % 3];
        // slice s has landed (own loads: at most the 4 of slice s + 1 still in
        // flight) and is visible from every wave; every wave is done with the
        // buffer of step s - 1, which slice s + 2 refills
        if (s + 1 < KS) __builtin_amdgcn_s_waitcnt(0x0074);   // vmcnt(4) lgkmcnt(0)
        else __builtin_amdgcn_s_waitcnt(0);
        wg_barrier();
        if (s + 2 < KS) stage_slice<KS>(wf, s + 2, wst[(G0 + s + 2) % 3], tid);
        __builtin_amdgcn_sched_barrier(0);   // issue the refill before this step's MFMAs
        // A of tile m + 1 is read while the MFMAs of tile m run
        bf16x8 A = buf[lane];
#pragma unroll
        for (int m = 0; m < 16; m++) {
            const bf16x8 An = buf[((m + 1) & 15) * 64 + lane];
#pragma unroll
            for (int t = 0; t < NT; t++)
                acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, bfrag(s, t), acc[m][t], 0, 0, 0);
            A = An;
        }
    }
}

constexpr int FEAT_LD = 72;   // LDS row of one parcel's 64 padded features (+8: bank spread)
constexpr int INFER_NT = 2;   // parcel tiles of 16 per wave (two waves per SIMD)
constexpr int INFER_WP = 16 * INFER_NT;   // parcels per wave

}  // namespace

__global__ void build_gelu_table_kernel(float2* tab) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= GTAB_N) return;
    const float p0 = gtab_phi(k), p1 = gtab_phi(k + 1);
    tab[k] = make_float2(p0, p1 - p0);
}

hipError_t launch_gelu_table(float2* tab, hipStream_t s) {
    hipLaunchKernelGGL(build_gelu_table_kernel, dim3(GTAB_N / 256), dim3(256), 0, s, tab);
    return hipGetLastError();
}

// one workgroup = 4 waves x INFER_WP parcels; every wave runs the network on
// its own parcels, the workgroup shares the staged weights and the GELU table
__global__ void __launch_bounds__(256, 4 / INFER_NT) surrogate_infer_kernel(SurrogateInferArgs a) {
    constexpr int NT = INFER_NT, WP = INFER_WP;
    __shared__ __bf16 fs[4][WP][FEAT_LD];
    __shared__ bf16x8 wst0[WSLICE], wst1[WSLICE], wst2[WSLICE];
    bf16x8* const wst[3] = {wst0, wst1, wst2};
    __shared__ float2 gtab[GTAB_N];
    __shared__ float sS[4 * WP][6];   // Sachs term of each parcel's Cartesian Eij
    __shared__ int sfl[4 * WP];       // feature-phase flags (envelope, audit)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    {   // GELU table: 8 KB, two 16-byte loads per thread
        const f32x4* src = (const f32x4*)a.gtab;
        f32x4* dst = (f32x4*)gtab;
        dst[tid] = src[tid];
        dst[tid + 256] = src[tid + 256];
    }
    const int64_t base = ((int64_t)blockIdx.x * 4 + wave) * WP;
    const int64_t P = a.fa.P;
    if (tid < 4 * WP) {   // features of the workgroup's parcels, one lane each; past P: zeros
        const int64_t p = (int64_t)blockIdx.x * 4 * WP + tid;
        const int fw = tid / WP, fp = tid % WP;
        float f[36];
        int fl = 0;
        if (p < P) {
            fl = parcel_features(a.fa, p, f, &a.mix, sS[tid]);
#pragma unroll
            for (int k = 0; k < 36; k++)
                fs[fw][fp][k] = (__bf16)((f[k] - a.norm[k]) * __builtin_amdgcn_rcpf(a.norm[36 + k]));
        } else {
#pragma unroll
            for (int k = 0; k < 36; k++) fs[fw][fp][k] = (__bf16)0.0f;
        }
#pragma unroll
        for (int k = 36; k < 64; k++) fs[fw][fp][k] = (__bf16)0.0f;
        sfl[tid] = fl;
        // flagged parcels (fallback / audit) join the work list: one atomic per wave
        if (a.idx) {
            const uint64_t m = __ballot(fl != 0);
            if (m) {
                const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                const int leader = __builtin_ctzll(m);
                int slot = 0;
                if (lane == leader) slot = atomicAdd(a.count, __popcll(m));
                slot = __shfl(slot, leader);
                if (fl) a.idx[slot + below] = (int)p;
            }
        }
    }
    // (the first barrier of mfma_layer orders these LDS writes before the reads)
    const int g = lane >> 4, c = lane & 15;
    f32x4 acc[16][NT];
    bf16x8 B[8][NT];
    // input layer, K = 64 (36 features + zero padding), natural K order
    acc_bias<NT>(a.bias[0], g, acc);
    mfma_layer<2, 0, NT>((const bf16x8*)a.wf[0],
                      [&](int s, int t) { return *(const bf16x8*)&fs[wave][16 * t + c][32 * s + 8 * g]; }, acc,
                      wst, tid, lane);
    gelu_repack<NT>(gtab, acc, B);
    acc_bias<NT>(a.bias[1], g, acc);
    mfma_layer<8, 2, NT>((const bf16x8*)a.wf[1], [&](int s, int t) { return B[s][t]; }, acc, wst, tid, lane);
    gelu_repack<NT>(gtab, acc, B);
    acc_bias<NT>(a.bias[2], g, acc);
    mfma_layer<8, 10, NT>((const bf16x8*)a.wf[2], [&](int s, int t) { return B[s][t]; }, acc, wst, tid, lane);
    gelu_repack<NT>(gtab, acc, B);
    // output layer: 16 rows (9 outputs + zero padding)
    f32x4 o[NT];
    {
        const f32x4 b = *(const f32x4*)(a.bias[3] + 4 * g);
#pragma unroll
        for (int t = 0; t < NT; t++) o[t] = b;
    }
    const bf16x8* wf3 = (const bf16x8*)a.wf[3];
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const bf16x8 A3 = wf3[s * 64 + lane];
#pragma unroll
        for (int t = 0; t < NT; t++) o[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A3, B[s][t], o[t], 0, 0, 0);
    }
    // outputs 4 g + i of parcel base + 16 t + c: de-normalise, exp.  The six
    // Eij outputs are the Taylor term T_ij; Eij = (1 - alpha) S_ij + alpha T_ij
    // with the Sachs term S from the feature phase, and the parcel leaves the
    // surrogate when alpha T_ij dT_ij > tol E_ij for any ij (or E is not a
    // positive finite number).  Lanes g = 0 and g = 1 hold the six Eij of a
    // parcel, g = 0 decides (the feature lanes' flags are in LDS: the parcel's
    // flags are written once, here).
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const int loc = wave * WP + 16 * t + c;
        const int64_t q = base + 16 * t + c;
        int bad = 0;
        if (q < P) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int oi = 4 * g + i;
                if (oi >= 9) continue;
                const float y = o[t][i] * a.norm[81 + oi] + a.norm[72 + oi];
                const double v = (double)expf(y);
                if (oi < 6) {
                    const double E = (1.0 - a.mix.alpha) * (double)sS[loc][oi] + a.mix.alpha * v;
                    a.Eij[q * 6 + oi] = E;
                    if (!(a.mix.alpha * v * a.mix.dT[oi] <= a.mix.tol * E) || !(E > 0.0) || !isfinite(E)) bad = 1;
                } else {
                    a.Vi[q * 3 + oi - 6] = v;
                }
            }
        }
        bad |= __shfl_xor(bad, 16);
        const bool own = g == 0 && q < P;
        int fl = own ? sfl[loc] : 0;
        const bool over = own && bad;
        const bool join = over && fl == 0;   // not on the work list yet
        if (over) fl |= SUR_BOUND;
        if (a.idx) {
            const uint64_t m = __ballot(join);
            if (m) {
                const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                const int leader = __builtin_ctzll(m);
                int slot = 0;
                if (lane == leader) slot = atomicAdd(a.count, __popcll(m));
                slot = __shfl(slot, leader);
                if (join) a.idx[slot + below] = (int)q;
            }
        }
        if (own && a.fa.flags) a.fa.flags[q] = fl;
    }
}

hipError_t launch_surrogate_infer(const SurrogateInferArgs& a, hipStream_t s) {
    if (a.fa.P == 0) return hipSuccess;
    if (!a.gtab) return hipErrorInvalidValue;
    hipLaunchKernelGGL(surrogate_infer_kernel, dim3(nb(a.fa.P, 4 * INFER_WP)), dim3(256), 0, s, a);
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

hipError_t launch_merge(const int* count, int64_t cap, const int* idx, const int* flags, const double* Ep,
                        const double* Vp, double* Eij, double* Vi, int* status, double* errE, double* errV,
                        hipStream_t s) {
    if (cap == 0) return hipSuccess;
    const unsigned grid = nb(cap, 256) < 1024u ? nb(cap, 256) : 1024u;
    hipLaunchKernelGGL(merge_kernel, dim3(grid), dim3(256), 0, s, count, cap, idx, flags, Ep, Vp, Eij, Vi, status,
                       errE, errV);
    return hipGetLastError();
}

hipError_t launch_audit_stats(const int* count, int64_t cap, const double* errE, const double* errV, AuditSums* out,
                              hipStream_t s) {
    hipLaunchKernelGGL(audit_stats_kernel, dim3(1), dim3(256), 0, s, count, cap, errE, errV, out);
    return hipGetLastError();
}

hipError_t launch_reset_count(int* count, hipStream_t s) {
    hipLaunchKernelGGL(reset_count_kernel, dim3(1), dim3(1), 0, s, count);
    return hipGetLastError();
}

}  // namespace sfg
