// Surrogate homogenization: feature extraction, training targets, output
// decoding, envelope / audit selection and the merge of the plain fallback.
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

__global__ void features_kernel(SurrogateFeatArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.P) return;
    double xq[15];
#pragma unroll
    for (int q = 0; q < 15; q++) xq[q] = a.x[(int64_t)q * a.P + p];
    const double x0 = xq[0];
    float* f = a.feat + p * 36;
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
        int fl = 0;
        for (int i = 0; i < 3; i++)
            if (!(lam[i] >= a.env_lo[i] && lam[i] <= a.env_hi[i])) fl |= SUR_FALLBACK;
        if (!(fl & SUR_FALLBACK) && unit_hash(a.seed ^ (uint64_t)p) < a.audit_frac) fl |= SUR_AUDIT;
        a.flags[p] = fl;
    }
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

// network outputs (raw units) -> Eij, sorted velocities
__global__ void decode_kernel(int64_t n, const float* y, double* Eij, double* Vi) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    for (int i = 0; i < 6; i++) Eij[p * 6 + i] = exp((double)y[p * 9 + i]);
    for (int i = 0; i < 3; i++) Vi[p * 3 + i] = exp((double)y[p * 9 + 6 + i]);
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

hipError_t launch_decode(int64_t n, const float* y, double* Eij, double* Vi, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(decode_kernel, dim3(nb(n, 256)), dim3(256), 0, s, n, y, Eij, Vi);
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
