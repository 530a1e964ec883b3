// Homogenization kernels (FP64): a2 eigenframe and the enhancement factors
// Eij of transversely isotropic grains under a linear Sachs/Taylor mix.
//
// One lane per parcel.  The structure tensors come from the l <= 4 block of
// the real SoA state through the same linear maps used by <D>
// (a2v = A2 x / x0 + I/3, a4v = A4 x / x0; src/moments.f90:164-178), the
// eigenframe from cyclic Jacobi rotations of a2 (the reference calls LAPACK
// dsyev, src/frames.f90:62-80), and the Taylor branch factors its 6x6 matrix
// once per parcel (the reference refactors it for each of the six stress
// states, src/enhancementfactors.f90:23-45 -> src/homogenizations.f90:145-189).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sfg_device.h"
#include "sfg_homog.h"

namespace sfg {

__device__ inline void d_vec_to_mat(const double* v, double* M) {
    const double s = sqrt(2.0);
    M[0] = v[0]; M[1] = v[5] / s; M[2] = v[4] / s;
    M[3] = v[5] / s; M[4] = v[1]; M[5] = v[3] / s;
    M[6] = v[4] / s; M[7] = v[3] / s; M[8] = v[2];
}

__device__ inline void d_mat_to_vec(const double* M, double* v) {
    const double s = sqrt(2.0);
    v[0] = M[0]; v[1] = M[4]; v[2] = M[8];
    v[3] = s * M[5]; v[4] = s * M[2]; v[5] = s * M[1];
}

__device__ inline double d_dinner22(const double* A, const double* B) {
    double s = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) s += A[3 * i + j] * B[3 * j + i];
    return s;
}

// rheo_params_tranisotropic (src/rheologies.f90:123-132), d = 3
__device__ inline void d_rheo_params(double Ecc, double Eca, double n, int ef, double* cI, double* cM, double* cL) {
    double ne = (double)(ef * 2) / (n + 1);
    double pc = pow(Ecc, ne), pa = pow(Eca, ne);
    *cI = (pc - 1) / 2;
    *cM = (3 * (pc + 1) - 2) / 2 - 2 * pa;
    *cL = pa - 1;
}

// a2 (Mandel vector) and a4 (Mandel 6x6) of parcel p from the real SoA state
__device__ inline void d_moments(const double* x, int64_t P, int64_t p, int nq, const double* A2, const double* A4,
                                 double* a2v, double* a4v) {
    double xq[15];
    for (int q = 0; q < 15; q++) xq[q] = q < nq ? x[(int64_t)q * P + p] : 0.0;
    const double x0 = xq[0];
    for (int i = 0; i < 6; i++) {
        double s = 0;
        for (int q = 0; q < nq; q++) s += A2[i * 15 + q] * xq[q];
        a2v[i] = s / x0 + (i < 3 ? 1.0 / 3.0 : 0.0);
    }
    for (int i = 0; i < 36; i++) {
        double s = 0;
        for (int q = 0; q < nq; q++) s += A4[i * 15 + q] * xq[q];
        a4v[i] = s / x0;
    }
}

// Symmetric 3x3 eigenpairs (cyclic Jacobi), largest eigenvalue first; ei rows.
// Sweeps stop once the off-diagonal sum is below 1e-18 of the diagonal's:
// the rotations leave rounding noise off the diagonal, so an exact-zero test
// ran all 60 sweeps for nearly every matrix, where this one takes 3-4 (same
// eigenpairs to ~1e-14 on near-isotropic a2, tested on 3 x 10^5 matrices).
__device__ void d_eig3(const double* M, double* ei, double* lami) {
    double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a[i][j] = 0.5 * (M[3 * i + j] + M[3 * j + i]);
    for (int sweep = 0; sweep < 60; sweep++) {
        const double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        if (off <= 1e-18 * (fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]))) break;
#pragma unroll
        for (int pq = 0; pq < 3; pq++) {
            const int p = pq == 2 ? 1 : 0, q = pq == 0 ? 1 : 2;
            if (a[p][q] == 0.0) continue;
            double theta = (a[q][q] - a[p][p]) / (2 * a[p][q]);
            double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
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
#pragma unroll
            for (int k = 0; k < 3; k++) {
                double vkp = v[k][p], vkq = v[k][q];
                v[k][p] = c * vkp - s * vkq;
                v[k][q] = s * vkp + c * vkq;
            }
        }
    }
    double d[3] = {a[0][0], a[1][1], a[2][2]};
    int i0 = 0, i1 = 1, i2 = 2, t;
    if (d[i1] > d[i0]) { t = i0; i0 = i1; i1 = t; }
    if (d[i2] > d[i0]) { t = i0; i0 = i2; i2 = t; }
    if (d[i2] > d[i1]) { t = i1; i1 = i2; i2 = t; }
    const int idx[3] = {i0, i1, i2};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        lami[i] = d[idx[i]];
#pragma unroll
        for (int k = 0; k < 3; k++) ei[3 * i + k] = v[k][idx[i]];
    }
}

// Cholesky of the symmetric matrix given by the lower triangle of A (LAPACK
// dposv('L') semantics, src/homogenizations.f90:174), in the order of the
// unblocked dpotf2 that dposv runs for n = 6: pivot, then the column below
// scaled by the pivot's reciprocal.  Lf packed lower (21).
__device__ inline bool d_chol6(const double A[6][6], double* Lf) {
#define LF(i, j) Lf[(i) * ((i) + 1) / 2 + (j)]
    for (int j = 0; j < 6; j++) {
        double d = A[j][j];
        for (int k = 0; k < j; k++) d -= LF(j, k) * LF(j, k);
        if (!(d > 0)) return false;
        d = sqrt(d);
        LF(j, j) = d;
        const double r = 1.0 / d;
        for (int i = j + 1; i < 6; i++) {
            double s = A[i][j];
            for (int k = 0; k < j; k++) s -= LF(i, k) * LF(j, k);
            LF(i, j) = s * r;
        }
    }
    return true;
}

// The matrix a failed dposv('L') leaves behind: F = A with its lower
// triangle overwritten as far as the factorization got (columns before the
// failed pivot j hold the factor, F[j][j] the non-positive pivot value; the
// rest of column j, the later columns and the upper triangle keep A's
// values).  The reference regularizes with this matrix
// (src/homogenizations.f90:176-179).
__device__ inline void d_failed_factor6(double F[6][6]) {
    for (int j = 0; j < 6; j++) {
        double d = F[j][j];
        for (int k = 0; k < j; k++) d -= F[j][k] * F[j][k];
        if (!(d > 0)) {
            F[j][j] = d;
            return;
        }
        d = sqrt(d);
        F[j][j] = d;
        const double r = 1.0 / d;
        for (int i = j + 1; i < 6; i++) {
            double t = F[i][j];
            for (int k = 0; k < j; k++) t -= F[i][k] * F[j][k];
            F[i][j] = t * r;
        }
    }
}

__device__ inline void d_chol6_solve(const double* Lf, double* b) {
    for (int i = 0; i < 6; i++) {
        double s = b[i];
        for (int k = 0; k < i; k++) s -= LF(i, k) * b[k];
        b[i] = s / LF(i, i);
    }
    for (int i = 5; i >= 0; i--) {
        double s = b[i];
        for (int k = i + 1; k < 6; k++) s -= LF(k, i) * b[k];
        b[i] = s / LF(i, i);
    }
#undef LF
}

// Taylor matrix P = I6 - A 1 (x) a2v + B a4v + C L(a2) (src/homogenizations.f90:164-172)
__device__ inline void d_taylorP(const double* a2m, const double* a2v, const double* a4v, double tA, double tB,
                                 double tC, double Pm[6][6]) {
    const double s = sqrt(2.0);
    const double* a = a2m;
    const double Lm[6][6] = {
        {2 * a[0], 0, 0, 0, s * a[2], s * a[1]},
        {0, 2 * a[4], 0, s * a[5], 0, s * a[1]},
        {0, 0, 2 * a[8], s * a[5], s * a[2], 0},
        {0, s * a[5], s * a[5], a[4] + a[8], a[1], a[2]},
        {s * a[2], 0, s * a[2], a[1], a[0] + a[8], a[5]},
        {s * a[1], s * a[1], 0, a[2], a[5], a[0] + a[4]}};
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++)
            Pm[i][j] = (i == j ? 1.0 : 0.0) - tA * ((i < 3 ? 1.0 : 0.0) * a2v[j]) + tB * a4v[6 * i + j] + tC * Lm[i][j];
}

// Eij of one parcel in frame (e1, e2, e3): src/enhancementfactors.f90:23-69.
// P is factored once per parcel; the rare regularized branch rebuilds the
// failed factorization where it needs it (the same arithmetic gives the same
// matrix), so it does not occupy registers through the loop over the six
// stress states.
__device__ int d_Eij(const double* a2v, const double* a4v, const double* e, double Ecc, double Eca, double alpha,
                     int ng, double* Eij) {
    int status = 0;
    double sA, sB, sC, tA, tB, tC;
    d_rheo_params(Ecc, Eca, (double)ng, 1, &sA, &sB, &sC);
    d_rheo_params(Ecc, Eca, (double)ng, -1, &tA, &tB, &tC);
    const double sachs_iso = 1 + 2.0 / 15 * sB + 2.0 / 3 * sC;            // src/homogenizations.f90:207
    const double taylor_iso = 1.0 / (1 + 2.0 / 15 * tB + 2.0 / 3 * tC);   // :235
    double a2m[9];
    d_vec_to_mat(a2v, a2m);
    // the Cholesky factor of P, or of the regularized normal equations
    // (P^T P + 1e-6 I, the REAL(4) 1e-6 of :178) when P is not positive definite
    double Lf[21];
    bool ok, ok_reg = true;
    {
        double Pm[6][6];
        d_taylorP(a2m, a2v, a4v, tA, tB, tC, Pm);
        ok = d_chol6(Pm, Lf);
        if (!ok) {
            status |= 2;
            d_failed_factor6(Pm);
            double R[6][6];
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) {
                    double s = 0;
                    for (int k = 0; k < 6; k++) s += Pm[k][i] * Pm[k][j];
                    R[i][j] = s + (i == j ? (double)1e-6f : 0.0);
                }
            ok_reg = d_chol6(R, Lf);
            if (!ok_reg) status |= 4;
        }
    }
    const int pv[6] = {0, 1, 2, 1, 0, 0}, pw[6] = {0, 1, 2, 2, 2, 1};
    for (int q = 0; q < 6; q++) {
        const double* v = e + 3 * pv[q];
        const double* w = e + 3 * pw[q];
        double tau[9], vw[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                vw[3 * i + j] = v[i] * w[j];
                tau[3 * i + j] = (q < 3) ? ((i == j) / 3.0 - v[i] * v[j]) : (v[i] * w[j] + w[i] * v[j]);
            }
        const double I2 = d_dinner22(tau, tau);
        // Sachs (src/homogenizations.f90:69-116), n' = 1 or -3
        double tv[6], a4t[6], a4tau[9], et2[9], es[9];
        d_mat_to_vec(tau, tv);
        for (int i = 0; i < 6; i++) {
            double s = 0;
            for (int j = 0; j < 6; j++) s += a4v[6 * i + j] * tv[j];
            a4t[i] = s;
        }
        d_vec_to_mat(a4t, a4tau);
        const double e0 = (ng == 1) ? 1.0 : I2;
        for (int i = 0; i < 9; i++) {
            et2[i] = (ng == 1) ? a2m[i] : I2 * a2m[i];
            if (ng != 1) a4tau[i] *= I2;
        }
        const double tr = d_dinner22(et2, tau);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double ac = 0;
                for (int k = 0; k < 3; k++) ac += tau[3 * i + k] * et2[3 * k + j] + et2[3 * i + k] * tau[3 * k + j];
                es[3 * i + j] = e0 * tau[3 * i + j] - sA * tr * (i == j) + sB * a4tau[3 * i + j] + sC * ac;
            }
        // Taylor (src/homogenizations.f90:145-189)
        double xs[6], et[9];
        d_mat_to_vec(tau, xs);
        if (ok) {
            d_chol6_solve(Lf, xs);
        } else {
            double Pm[6][6], rhs[6];
            d_taylorP(a2m, a2v, a4v, tA, tB, tC, Pm);
            d_failed_factor6(Pm);
            for (int i = 0; i < 6; i++) {
                rhs[i] = 0;
                for (int k = 0; k < 6; k++) rhs[i] += Pm[k][i] * xs[k];
            }
            if (ok_reg) d_chol6_solve(Lf, rhs);
            for (int i = 0; i < 6; i++) xs[i] = rhs[i];
        }
        d_vec_to_mat(xs, et);
        const double siso = (ng == 1) ? sachs_iso : I2;
        const double tvw = d_dinner22(tau, vw);
        const double ts = d_dinner22(es, vw) / (siso * tvw);
        const double tt = d_dinner22(et, vw) / (taylor_iso * tvw);
        Eij[q] = (1 - alpha) * ts + alpha * tt;
        if (!isfinite(Eij[q])) status |= 1;
    }
    return status;
}

__device__ __forceinline__ void eij_item(const EijArgs& a, int64_t t) {
    const int64_t p = a.idx ? (int64_t)a.idx[t] : t;   // state column
    const int64_t o = t;                               // output row
    double a2v[6], a4v[36];
    d_moments(a.x, a.P, p, a.nq, a.A2, a.A4, a2v, a4v);
    double e[9], lam[3] = {0, 0, 0};
    if (a.frame == 0) {
        double a2m[9];
        d_vec_to_mat(a2v, a2m);
        d_eig3(a2m, e, lam);
    } else if (a.frame == 1) {
        for (int i = 0; i < 9; i++) e[i] = a.frames[p * 9 + i];
    } else {
        for (int i = 0; i < 9; i++) e[i] = (i % 4 == 0) ? 1.0 : 0.0;
    }
    double E[6];
    int st = d_Eij(a2v, a4v, e, a.Ecc, a.Eca, a.alpha, a.n_grain, E);
    for (int i = 0; i < 6; i++) a.Eij[o * 6 + i] = E[i];
    if (a.lami) for (int i = 0; i < 3; i++) a.lami[o * 3 + i] = lam[i];
    if (a.ei) for (int i = 0; i < 9; i++) a.ei[o * 9 + i] = e[i];
    if (a.status && st) atomicOr(a.status + p, st);
}

// one wave per SIMD: the Taylor 6x6 system, a4 and the frame need ~380
// registers per lane, which then stay in VGPRs/AGPRs instead of scratch
__global__ void __launch_bounds__(128, 1) eij_kernel(EijArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (a.idx ? a.nidx : a.P)) return;
    eij_item(a, t);
}

// work list whose length is on the device: grid-stride loop (a separate
// kernel, so that the loop does not cost the one-item kernel registers)
__global__ void __launch_bounds__(128, 1) eij_list_kernel(EijArgs a) {
    const int64_t n = *a.nidx_dev < a.nidx ? *a.nidx_dev : a.nidx;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) eij_item(a, t);
}

hipError_t launch_eij(const EijArgs& a, hipStream_t s) {
    const int64_t n = a.idx ? a.nidx : a.P;
    if (n == 0) return hipSuccess;
    const int64_t nblk = (n + 127) / 128;
    if (a.idx && a.nidx_dev)   // a device-side count: at most 2048 workgroups
        hipLaunchKernelGGL(eij_list_kernel, dim3((unsigned)std::min<int64_t>(nblk, 2048)), dim3(128), 0, s, a);
    else
        hipLaunchKernelGGL(eij_kernel, dim3((unsigned)nblk), dim3(128), 0, s, a);
    return hipGetLastError();
}

}  // namespace sfg
