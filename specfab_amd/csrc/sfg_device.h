// Device-side helpers shared by the HIP kernels: a minimal complex type and
// the per-parcel weight formulas of the reference, evaluated in FP64 with the
// reference's REAL(4) literals where it has them.
#pragma once
#ifdef SFG_HOST_EMUL
#include "sfg_emul.h"
#else
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

namespace sfg {

struct cplx {
    double re, im;
};
__device__ __forceinline__ cplx cmk(double r, double i) { return cplx{r, i}; }
__device__ __forceinline__ cplx operator+(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cplx operator-(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cplx operator-(cplx a) { return {-a.re, -a.im}; }
__device__ __forceinline__ cplx operator*(cplx a, cplx b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cplx operator*(double s, cplx a) { return {s * a.re, s * a.im}; }
__device__ __forceinline__ cplx operator*(cplx a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ cplx operator/(cplx a, double s) { return {a.re / s, a.im / s}; }

#define SFG_PI 3.141592653589793

// quad_rr (src/dynamics.f90:563-580): l=2 coefficients of M:rr, q[0..4] = m=-2..2
__device__ inline void d_quad_rr(const double* M, cplx* q) {
    const double fsq = sqrt(2 * SFG_PI / 15);
    const double c23 = (double)(2.0f / 3.0f);   // REAL(4) literal 2./3
    double xx = M[0], yy = M[4], zz = M[8], xy = M[1], xz = M[2], yz = M[5];
    q[0] = fsq * cmk(xx - yy, 2 * xy);
    q[1] = (2 * fsq) * cmk(xz, yz);
    q[2] = cmk(-(c23 * sqrt(SFG_PI / 5) * (xx + yy - 2 * zz)), 0.0);
    q[3] = (-2 * fsq) * cmk(xz, -yz);
    q[4] = fsq * cmk(xx - yy, -2 * xy);
}

// quad_tp (src/dynamics.f90:582-593): l=1 coefficients of M:tp, q[0..2] = m=-1..1
__device__ inline void d_quad_tp(const double* M, cplx* q) {
    const double fsq1 = sqrt(2 * SFG_PI / 3);
    const double s2 = (double)sqrtf(2.0f);
    double xy = M[1], xz = M[2], yz = M[5];
    q[0] = fsq1 * cmk(yz, -xz);
    q[1] = cmk(fsq1 * (s2 * xy), 0.0);
    q[2] = fsq1 * cmk(-yz, -xz);
}

// DDRX weights g(1:15) * k*5/(tau:tau): src/include/ddrx-coupling-weights.f90
// and src/dynamics.f90:291-293.
__device__ inline void d_ddrx_weights(const double* tau, cplx* g) {
    cplx q[5];
    d_quad_rr(tau, q);
    cplx qm2 = q[0], qm1 = q[1], q0 = q[2], qp1 = q[3], qp2 = q[4];
    const double k = (3 * sqrt(5 / SFG_PI)) / 28.0;
    const double S5 = sqrtf(5.0f), S15 = sqrtf(1.5f), S6 = sqrtf(6.0f), S2 = sqrtf(2.0f), S3 = sqrtf(3.0f);
    const double m2S14 = (float)(-2.0f * sqrtf(14.0f)), m4S7 = (float)(-4.0f * sqrtf(7.0f));
    const double t3S5 = (float)(3.0f * sqrtf(5.0f));
    g[0] = (7.0 * (q0 * q0 + (-2.0) * qm1 * qp1 + 2.0 * qm2 * qp2)) / S5;
    g[1] = S15 * (qm1 * qm1) + (-2.0) * q0 * qm2;
    g[2] = q0 * qm1 + ((-1.0) * S6) * qp1 * qm2;
    g[3] = q0 * q0 + (-1.0) * qm1 * qp1 + (-2.0) * qm2 * qp2;
    g[4] = q0 * qp1 + ((-1.0) * S6) * qm1 * qp2;
    g[5] = S15 * (qp1 * qp1) + (-2.0) * q0 * qp2;
    g[6] = (m2S14 * (qm2 * qm2)) / 3.0;
    g[7] = (m4S7 * qm1 * qm2) / 3.0;
    g[8] = (-4.0 * (S2 * (qm1 * qm1) + S3 * q0 * qm2)) / 3.0;
    g[9] = (-4.0 * (S6 * q0 * qm1 + qp1 * qm2)) / 3.0;
    g[10] = (-4.0 * (3.0 * (q0 * q0) + 4.0 * qm1 * qp1 + qm2 * qp2)) / t3S5;
    g[11] = (-4.0 * (S6 * q0 * qp1 + qm1 * qp2)) / 3.0;
    g[12] = (-4.0 * (S2 * (qp1 * qp1) + S3 * q0 * qp2)) / 3.0;
    g[13] = (m4S7 * qp1 * qp2) / 3.0;
    g[14] = (m2S14 * (qp2 * qp2)) / 3.0;
    double tt = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) tt += tau[3 * i + j] * tau[3 * j + i];
    for (int m = 0; m < 15; m++) g[m] = ((k * g[m]) * 5.0) / tt;
}

// LROT weights (g0, gz, gn, gp)(k = 0..5): src/dynamics.f90:67-91
__device__ inline void d_lrot_weights(const double* eps, const double* omg, double iota, double zeta, cplx* g0,
                                      cplx* gz, cplx* gn, cplx* gp, cplx* qe_out, cplx* qo_out) {
    double epssq[9], arg[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += eps[3 * i + k] * eps[3 * k + j];
            epssq[3 * i + j] = s;
        }
    double zetanorm = zeta / sqrt(epssq[0] + epssq[4] + epssq[8]);   // NaN for eps = 0 (as the reference)
    for (int i = 0; i < 9; i++) arg[i] = iota * eps[i] + zetanorm * epssq[i];
    cplx qe[5], qo[3];
    d_quad_rr(arg, qe);
    d_quad_tp(omg, qo);
    for (int i = 0; i < 5; i++) qe_out[i] = qe[i];
    for (int i = 0; i < 3; i++) qo_out[i] = qo[i];
    if (!g0) return;
    const double s3 = (double)sqrtf(3.0f), s6 = (double)sqrtf(6.0f);
    const double c56 = (double)sqrtf(5.0f / 6.0f), c23 = (double)sqrtf(2.0f / 3.0f), c32 = (double)sqrtf(3.0f / 2.0f);
    const cplx I = cmk(0, 1);
    cplx gz0 = -((s3 * I) * qo[1]);
    cplx gn0 = -(((6.0 * I) / s6) * qo[0]);
    cplx gp0 = ((6.0 * I) / s6) * qo[2];
    g0[0] = cmk(0, 0);
    for (int k = 1; k < 6; k++) g0[k] = 3.0 * qe[k - 1];
    gz[0] = gz0; gz[1] = -qe[0]; gz[2] = cmk(0, 0); gz[3] = cmk(0, 0); gz[4] = cmk(0, 0); gz[5] = qe[4];
    gn[0] = gn0 + c56 * qe[1]; gn[1] = cmk(0, 0); gn[2] = qe[0]; gn[3] = c23 * qe[1]; gn[4] = c32 * qe[2];
    gn[5] = 2.0 * qe[3];
    gp[0] = gp0 + c56 * qe[3]; gp[1] = 2.0 * qe[1]; gp[2] = c32 * qe[2]; gp[3] = c23 * qe[3]; gp[4] = qe[4];
    gp[5] = cmk(0, 0);
}

__device__ __forceinline__ void sym_skew(const double* G, double* D, double* W) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            D[3 * i + j] = (G[3 * i + j] + G[3 * j + i]) / 2;
            W[3 * i + j] = (G[3 * i + j] - G[3 * j + i]) / 2;
        }
}

}  // namespace sfg
