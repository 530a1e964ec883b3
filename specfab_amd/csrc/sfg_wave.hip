// Elastic and electromagnetic phase velocities of transversely isotropic
// grains, and the full fourth-order structure tensor (FP64).
//
// One lane per (parcel, direction) pair, parcels fastest so that a wavefront
// reads one coalesced row of the SoA state per coefficient.  Per lane
// (src/wavepropagation.f90:121-144, 217-261):
//   1. rotate the l <= 4 block by the direction's real rotation blocks
//      (rotate_nlm4 at (-theta, -phi), tabulated on the host per direction);
//   2. structure tensors of the rotated state through linear maps: a2 and
//      Mandel a4 (REAL(8) bodies) for the Voigt branch, the index-sorted full
//      a4 (REAL(4) body) for the Reuss branch;
//   3. Reuss: the sym-sym block of the 9x9 compliance of
//      src/include/P_Reuss.f90 is the Mandel 6x6
//        P = k1 1(x)1 + k2 I + k3 (1(x)a2 + a2(x)1) + k4 a4 + k5 {a2, .};
//      antisymmetric stresses decouple from it, so the 9x9 inverse the
//      reference forms (matinv, :326-361) is replaced by a pivoted 6x6 LU
//      with the three right-hand sides sym(e_j (x) z) that Qnorm reads;
//   4. Voigt acoustic tensor (:170-187), mix (1-alpha) R + alpha V;
//   5. Christoffel roots by the reference's complex Cardano formula with its
//      REAL(4) exponents (Qnorm_to_vi, :267-324).
#include <algorithm>

#include "sfg_device.h"
#include "generated/sfg_l4maps.h"
#include "sfg_a4slots.h"
#include "sfg_wave.h"

namespace sfg {

namespace {

// Mandel pair I -> index-sorted a4 entry of (pair I, pair J)
__constant__ int kMandelA4[6][6] = {{0, 3, 5, 4, 2, 1},  {3, 10, 12, 11, 7, 6}, {5, 12, 14, 13, 9, 8},
                                    {4, 11, 13, 12, 8, 7}, {2, 7, 9, 8, 5, 4},   {1, 6, 8, 7, 4, 3}};

__device__ inline cplx zdiv(cplx a, cplx b) {
    double d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
__device__ inline cplx zlog(cplx z) { return {log(hypot(z.re, z.im)), atan2(z.im, z.re)}; }
__device__ inline cplx zexp(cplx z) {
    double s, c;
    sincos(z.im, &s, &c);
    double e = exp(z.re);
    return {e * c, e * s};
}
// principal z**w = exp(w log z)
__device__ inline cplx zpow(cplx z, cplx w) {
    if (z.re == 0.0 && z.im == 0.0) return (w.re == 0.0 && w.im == 0.0) ? cplx{1.0, 0.0} : cplx{0.0, 0.0};
    return zexp(w * zlog(z));
}
__device__ inline cplx zpow(cplx z, double w) { return zpow(z, cplx{w, 0.0}); }
__device__ inline cplx zsqrt(cplx z) {
    double r = hypot(z.re, z.im);
    if (r == 0.0) return {0.0, z.im};
    if (z.re >= 0) {
        double t = sqrt(0.5 * (r + z.re));
        return {t, z.im / (2 * t)};
    }
    double t = sqrt(0.5 * (r - z.re));
    return {fabs(z.im) / (2 * t), copysign(t, z.im)};
}

// rotated l <= 4 block of parcel p (R2 then R4 row-major)
__device__ inline void rotated_block(const double* x, int64_t P, int64_t p, int nq, const double* R, double* xr) {
    double x4[15];
#pragma unroll
    for (int q = 0; q < 15; q++) x4[q] = q < nq ? x[(int64_t)q * P + p] : 0.0;
    xr[0] = x4[0];
#pragma unroll
    for (int i = 0; i < 5; i++) {
        double s = 0;
#pragma unroll
        for (int j = 0; j < 5; j++) s += R[5 * i + j] * x4[1 + j];
        xr[1 + i] = s;
    }
#pragma unroll
    for (int i = 0; i < 9; i++) {
        double s = 0;
#pragma unroll
        for (int j = 0; j < 9; j++) s += R[25 + 9 * i + j] * x4[6 + j];
        xr[6 + i] = s;
    }
}

__device__ inline double map_row(const double* A, int row, const double* xr, int nq) {
    double s = 0;
#pragma unroll
    for (int q = 0; q < 15; q++)
        if (q < nq) s += A[row * 15 + q] * xr[q];
    return s;
}

// Reuss acoustic tensor for k = z; returns false if the compliance is singular
__device__ bool qnorm_reuss(const double* a2v, const double* u, const double* k, double Q[9]) {
    const double r2 = sqrt(2.0);
    const double w[6] = {1, 1, 1, r2, r2, r2};
    const double one[6] = {1, 1, 1, 0, 0, 0};
    double a2[9];
    a2[0] = a2v[0]; a2[4] = a2v[1]; a2[8] = a2v[2];
    a2[5] = a2[7] = a2v[3] / r2; a2[2] = a2[6] = a2v[4] / r2; a2[1] = a2[3] = a2v[5] / r2;
    // Mandel anticommutator {a2, .} (src/homogenizations.f90:238-258)
    const double Lm[6][6] = {{2 * a2[0], 0, 0, 0, r2 * a2[2], r2 * a2[1]},
                             {0, 2 * a2[4], 0, r2 * a2[5], 0, r2 * a2[1]},
                             {0, 0, 2 * a2[8], r2 * a2[5], r2 * a2[2], 0},
                             {0, r2 * a2[5], r2 * a2[5], a2[4] + a2[8], a2[1], a2[2]},
                             {r2 * a2[2], 0, r2 * a2[2], a2[1], a2[0] + a2[8], a2[5]},
                             {r2 * a2[1], r2 * a2[1], 0, a2[2], a2[5], a2[0] + a2[4]}};
    double A[6][9];
#pragma unroll
    for (int I = 0; I < 6; I++) {
#pragma unroll
        for (int J = 0; J < 6; J++)
            A[I][J] = k[0] * one[I] * one[J] + (I == J ? k[1] : 0.0) + k[2] * (one[I] * a2v[J] + a2v[I] * one[J]) +
                      k[3] * (w[I] * w[J]) * u[kMandelA4[I][J]] + k[4] * Lm[I][J];
        // right-hand sides: Mandel vectors of sym(e_j (x) z), j = x, y, z
        A[I][6] = (I == 4) ? 1.0 / r2 : 0.0;
        A[I][7] = (I == 3) ? 1.0 / r2 : 0.0;
        A[I][8] = (I == 2) ? 1.0 : 0.0;
    }
    // LU with partial pivoting (first maximal pivot), forward elimination
#pragma unroll
    for (int c = 0; c < 6; c++) {
#pragma unroll
        for (int r = c + 1; r < 6; r++) {
            bool sw = fabs(A[r][c]) > fabs(A[c][c]);
#pragma unroll
            for (int j = c; j < 9; j++) {
                double t = A[c][j];
                A[c][j] = sw ? A[r][j] : t;
                A[r][j] = sw ? t : A[r][j];
            }
        }
        if (A[c][c] == 0.0) return false;
        const double inv = 1.0 / A[c][c];
#pragma unroll
        for (int r = c + 1; r < 6; r++) {
            const double f = A[r][c] * inv;
#pragma unroll
            for (int j = c + 1; j < 9; j++) A[r][j] -= f * A[c][j];
        }
    }
    // back substitution for the three columns
#pragma unroll
    for (int col = 6; col < 9; col++) {
#pragma unroll
        for (int r = 5; r >= 0; r--) {
            double s = A[r][col];
#pragma unroll
            for (int j = r + 1; j < 6; j++) s -= A[r][j] * A[j][col];
            A[r][col] = s / A[r][r];
        }
    }
    // Q_ij = stress_{i z} of the strain sym(e_j (x) z)
#pragma unroll
    for (int j = 0; j < 3; j++) {
        Q[0 * 3 + j] = A[4][6 + j] / r2;
        Q[1 * 3 + j] = A[3][6 + j] / r2;
        Q[2 * 3 + j] = A[2][6 + j];
    }
    return true;
}

// Voigt acoustic tensor for k = z (src/wavepropagation.f90:170-187);
// a4c = third Mandel column of a4 (a4 : zz)
__device__ void qnorm_voigt(const double* a2v, const double* a4c, const double* k, double Q[9]) {
    const double r2 = sqrt(2.0);
    double a2[9], a4z[9];
    a2[0] = a2v[0]; a2[4] = a2v[1]; a2[8] = a2v[2];
    a2[5] = a2[7] = a2v[3] / r2; a2[2] = a2[6] = a2v[4] / r2; a2[1] = a2[3] = a2v[5] / r2;
    a4z[0] = a4c[0]; a4z[4] = a4c[1]; a4z[8] = a4c[2];
    a4z[5] = a4z[7] = a4c[3] / r2; a4z[2] = a4z[6] = a4c[4] / r2; a4z[1] = a4z[3] = a4c[5] / r2;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const double id = (i == j) ? 1.0 : 0.0, zz = (i == 2 && j == 2) ? 1.0 : 0.0;
            const double KM = (i == 2 ? a2[3 * 2 + j] : 0.0) + (j == 2 ? a2[3 * i + 2] : 0.0);
            Q[3 * i + j] = k[1] / 2 * id + (k[0] + k[1] / 2) * zz + k[2] * KM + k[3] * a4z[3 * i + j] +
                           0.5 * k[4] * (a2[3 * i + j] + a2[8] * id + KM);
        }
}

// Qnorm_to_vi + detQ_polycoefs (src/wavepropagation.f90:267-324)
__device__ void qnorm_to_vi(const double* Q, double rho, double* vi) {
    const double w = rho;   // omega = 1
    const double I1 = Q[0] + Q[4] + Q[8];
    const double I2 = -Q[1] * Q[1] - Q[2] * Q[2] + Q[0] * Q[4] - Q[5] * Q[5] + Q[0] * Q[8] + Q[4] * Q[8];
    const double I3 = -Q[2] * Q[2] * Q[4] + 2 * Q[1] * Q[2] * Q[5] - Q[0] * Q[5] * Q[5] - Q[1] * Q[1] * Q[8] +
                      Q[0] * Q[4] * Q[8];
    const double r0 = -w * w * w, r2 = I1 * w * w, r4 = -(I2 * w), r6 = I3;
    // REAL(4) literals 0.6666666666666666, 0.3333333333333333, 1.3333333333333333
    const double e23 = (double)0.6666666666666666f, e13 = (double)0.3333333333333333f;
    const double e43 = (double)1.3333333333333333f;
    const cplx a{(3.0 * r6 * r2 - r4 * r4) / (3.0 * r6 * r6), 0.0};
    const cplx b{(2.0 * r4 * r4 * r4 - 9.0 * r6 * r4 * r2 + 27.0 * r6 * r6 * r0) / (27.0 * r6 * r6 * r6), 0.0};
    const cplx root = zsqrt(12.0 * zpow(a, 3.0) + 81.0 * zpow(b, 2.0));
    const cplx base = (-9.0) * b + root;
    const cplx denom = zpow(cplx{6.0, 0.0}, e23) * zpow(base, e13);
    const cplx D = zpow(base, e23);
    const cplx sh{-r4 / (3.0 * r6), 0.0};
    const cplx c3 = zpow(cplx{3.0, 0.0}, e13), c2 = zpow(cplx{2.0, 0.0}, e13);
    const cplx x1 = sh + zdiv((-2.0) * c3 * a + c2 * D, denom);
    const cplx x2 = sh + zdiv(zpow(cplx{-1.0, 0.0}, e13) * (2.0 * c3 * a + zpow(cplx{-2.0, 0.0}, e13) * D), denom);
    const cplx x3 = sh + zdiv(zpow(cplx{-1.0, 0.0}, e43) * (2.0 * zpow(cplx{-3.0, 0.0}, e13) * a + c2 * D), denom);
    const cplx q1 = zsqrt(x3), q2 = zsqrt(x1), q3 = zsqrt(x2);
    vi[0] = 1.0 / hypot(q1.re, q1.im);
    vi[1] = 1.0 / hypot(q2.re, q2.im);
    vi[2] = 1.0 / hypot(q3.re, q3.im);
}

}  // namespace

__device__ __forceinline__ void vi_elastic_item(const ViElasticArgs& a, int64_t t, int64_t nwork) {
    const int64_t item = t % nwork;
    const int d = (int)(t / nwork);
    const int64_t p = a.idx ? (int64_t)a.idx[item] : item;   // state column
    double xr[15];
    rotated_block(a.x, a.P, p, a.nq, a.R + (int64_t)d * ROT4_STRIDE, xr);
    const double x0 = xr[0];
    double a2v[6];
#pragma unroll
    for (int i = 0; i < 6; i++) a2v[i] = map_row(a.A2, i, xr, 15) / x0 + (i < 3 ? 1.0 / 3.0 : 0.0);
    double Q[9];
#pragma unroll
    for (int i = 0; i < 9; i++) Q[i] = 0.0;
    bool ok = true;
    if (a.alpha < 1.0) {
        double u[15], QR[9];
#pragma unroll
        for (int e = 0; e < 15; e++) u[e] = map_row(a.A4F, e, xr, 15) / x0;
        ok = qnorm_reuss(a2v, u, a.kf, QR);
#pragma unroll
        for (int i = 0; i < 9; i++) Q[i] += (1 - a.alpha) * QR[i];
    }
    if (a.alpha > 0.0) {
        double a4c[6], QV[9];
#pragma unroll
        for (int I = 0; I < 6; I++) a4c[I] = map_row(a.A4, 6 * I + 2, xr, 15) / x0;
        qnorm_voigt(a2v, a4c, a.kr, QV);
#pragma unroll
        for (int i = 0; i < 9; i++) Q[i] += a.alpha * QV[i];
    }
    if (a.Qn) {
        double* q = a.Qn + (item * a.ndir + d) * 9;
#pragma unroll
        for (int i = 0; i < 9; i++) q[i] = ok ? Q[i] : __builtin_nan("");
    }
    double vi[3];
    if (ok) {
        qnorm_to_vi(Q, a.rho, vi);
    } else {
        vi[0] = vi[1] = vi[2] = __builtin_nan("");
        if (a.status) atomicOr(a.status + p, 8);
    }
    double* o = a.Vi + (item * a.ndir + d) * 3;
    o[0] = vi[0];
    o[1] = vi[1];
    o[2] = vi[2];
}

__global__ void __launch_bounds__(128) vi_elastic_kernel(ViElasticArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nwork = a.idx ? a.nidx : a.P;
    if (t < nwork * a.ndir) vi_elastic_item(a, t, nwork);
}

// work list whose length is on the device: grid-stride loop
__global__ void __launch_bounds__(128) vi_elastic_list_kernel(ViElasticArgs a) {
    const int64_t nwork = *a.nidx_dev < a.nidx ? *a.nidx_dev : a.nidx;
    const int64_t n = nwork * a.ndir;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) vi_elastic_item(a, t, nwork);
}

__global__ void __launch_bounds__(128) vi_em_kernel(ViEMArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.P * a.ndir) return;
    const int64_t p = t % a.P;
    const int d = (int)(t / a.P);
    const double* R = a.R + (int64_t)d * ROT4_STRIDE;
    double x2[6], xr[6];
#pragma unroll
    for (int q = 0; q < 6; q++) x2[q] = a.x[(int64_t)q * a.P + p];
    xr[0] = x2[0];
#pragma unroll
    for (int i = 0; i < 5; i++) {
        double s = 0;
#pragma unroll
        for (int j = 0; j < 5; j++) s += R[5 * i + j] * x2[1 + j];
        xr[1 + i] = s;
    }
    double a2v[6];
#pragma unroll
    for (int i = 0; i < 6; i++) a2v[i] = map_row(a.A2, i, xr, 6) / xr[0] + (i < 3 ? 1.0 / 3.0 : 0.0);
    const double r2 = sqrt(2.0);
    double A[9];
    A[0] = a2v[0]; A[4] = a2v[1]; A[8] = a2v[2];
    A[5] = a2v[3] / r2; A[2] = a2v[4] / r2; A[1] = a2v[5] / r2;
    // eps = (2 eps_t + eps_m) I/3 + (eps_m - eps_t)(a2 - I/3)
    const double dm = a.eps_m - a.eps_t, iso = (2 * a.eps_t + a.eps_m) / 3.0;
    const double e11 = iso + dm * (A[0] - 1.0 / 3.0), e22 = iso + dm * (A[4] - 1.0 / 3.0);
    const double e33 = iso + dm * (A[8] - 1.0 / 3.0);
    const double e12 = dm * A[1], e13 = dm * A[2], e23 = dm * A[5];
    const double c = 299792458.0;
    const double w = (1.0 / c) * (1.0 / c) * a.mu;
    // q^2 roots of eps33 q^4 + bq q^2 + det(eps) w^2 = 0 (Mathematica export, :236-245)
    const double bq = (e13 * e13 + e23 * e23 - (e11 + e22) * e33) * w;
    const double t13 = e13 * e13 + e23 * e23;
    const double disc = (t13 * t13 - 2.0 * (e13 * e13 * (e11 - e22) + 4.0 * e12 * e13 * e23 + (e22 - e11) * e23 * e23) * e33 +
                         (4.0 * e12 * e12 + (e11 - e22) * (e11 - e22)) * e33 * e33) * (w * w);
    const double sd = sqrt(disc);
    const double qf = sqrt(-(1.0 / e33) * (bq + sd)) / sqrt(2.0);
    const double qs = sqrt((1.0 / e33) * (-bq + sd)) / sqrt(2.0);
    double* o = a.Vi + (p * a.ndir + d) * 2;
    o[0] = 1.0 / qf;
    o[1] = 1.0 / qs;
}

__global__ void __launch_bounds__(128) a4full_kernel(A4FullArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.P) return;
    double x4[15], u[15];
#pragma unroll
    for (int q = 0; q < 15; q++) x4[q] = q < a.nq ? a.x[(int64_t)q * a.P + p] : 0.0;
#pragma unroll
    for (int e = 0; e < 15; e++) u[e] = map_row(a.A4F, e, x4, 15) / x4[0];
    // entry (i,j,k,l) -> sorted entry; the reference body sets a4(3,2,1,2)
    // to a4(1,2,3,3) (src/include/ev_c4__body.f90), reproduced here
    double* o = a.a4 + p * 81;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++)
#pragma unroll
                for (int l = 0; l < 3; l++) {
                    int v0 = i, v1 = j, v2 = k, v3 = l, t;
                    if (v1 < v0) { t = v0; v0 = v1; v1 = t; }
                    if (v3 < v2) { t = v2; v2 = v3; v3 = t; }
                    if (v2 < v0) { t = v0; v0 = v2; v2 = t; }
                    if (v3 < v1) { t = v1; v1 = v3; v3 = t; }
                    if (v2 < v1) { t = v1; v1 = v2; v2 = t; }
                    // index of sorted (v0 <= v1 <= v2 <= v3) among the 15
                    int idx = 0;
                    for (int a0 = 0; a0 < 3; a0++)
                        for (int a1 = a0; a1 < 3; a1++)
                            for (int a2 = a1; a2 < 3; a2++)
                                for (int a3 = a2; a3 < 3; a3++) {
                                    if (a0 < v0 || (a0 == v0 && (a1 < v1 || (a1 == v1 && (a2 < v2 || (a2 == v2 && a3 < v3))))))
                                        idx++;
                                }
                    o[27 * i + 9 * j + 3 * k + l] = (i == 2 && j == 1 && k == 0 && l == 1) ? u[8] : u[idx];
                }
}


// ---------------------------------------------------------------------------
// Orthotropic grains.  Each slip-system axis distribution contributes its a2
// (the reference's a2 body) and its full a4 (the REAL(4) body, its (3,2,1,2)
// entry included), both from the compile-time maps of the l <= 4 block; the
// joint tensors follow from the two distributions' moments under the product
// density with the kernel sin^2 of the axes' angle, as the reference's
// generated bodies evaluate them (src/moments.f90:242-310; identified against
// them to their REAL(4) level):
//   a4_joint(b, n)      = (a2b (x) a2n - a4b : a4n) / (1 - a2b : a2n)       <b b n n>
//   a4_jointcross(b, n) = eps eps (a4n (x) a2b) / (1 - a2b : a2n)           <n n v v>
//   a2_orth(b, n)       = eps eps (a2b (x) a2n) / (1 - a2b : a2n)           <v v>
//   a4_orth(b, n)       = eps^4 (a4b (x) a4n) / (1 - 2 a2b:a2n + a4b::a4n)   <v v v v>
// with v = b x n and eps the Levi-Civita symbol.
//
// Sachs (src/homogenizations.f90:288-319): Q = sum_m ci_m M2_m + ci_m+3 H2_m
// with M2_m = (a4_jj + a4_kk - 2 sym2(J_m)) / 4 and H2_m = sym4(J_m).  Apart
// from the body a4's (3,2,1,2) entry, Q is symmetric within its index pairs
// (lk) and (ij), so it is held as a 6 x 6 pair matrix Q6 plus that entry's
// excess dq; the a4 terms combine into one symmetric tensor
// S = sum_d w_d a4_d, w_d = (ci_0 + ci_1 + ci_2 - ci_d) / 4, and each joint
// entry is added into Q6 where sym2 and sym4 place it (tabulated below).
// eps(tau) : vw is then a 6 x 6 contraction per stress state.

namespace {

// pair <-> indices: xx yy zz yz xz xy
__host__ __device__ constexpr int pa(int P) { return P < 3 ? P : (P == 3 ? 1 : 0); }
__host__ __device__ constexpr int pb(int P) { return P < 3 ? P : (P == 5 ? 1 : 2); }
__host__ __device__ constexpr int pidx(int a, int b) { return a == b ? a : 6 - a - b; }
__host__ __device__ constexpr double pw(int P) { return P < 3 ? 1.0 : 2.0; }   // orderings of a pair
__host__ __device__ constexpr int levi(int i, int j, int k) {
    return (i == j || j == k || i == k) ? 0 : (((j - i + 3) % 3 == 1) ? 1 : -1);
}
// sorted slot of the symmetric a4 entry (pair X, pair Y)
__host__ __device__ constexpr int sxy(int X, int Y) { return kA4Slots.sym[27 * pa(X) + 9 * pb(X) + 3 * pa(Y) + pb(Y)]; }

// Destinations of the pair entry J[X][Y] in sym4(J)(l,k,i,j) =
// (J(l,i,k,j) + J(i,k,l,j) + J(l,j,i,k) + J(j,k,i,l)) / 4, (l,k) = P,
// (i,j) = R: at most four pair entries (P, R), each receiving it c times
struct H2Dst {
    signed char P[36][4], R[36][4], c[36][4];
};
__host__ __device__ constexpr H2Dst make_h2dst() {
    H2Dst m{};
    for (int P = 0; P < 6; P++)
        for (int R = 0; R < 6; R++) {
            const int l = pa(P), k = pb(P), i = pa(R), j = pb(R);
            const int xy[4] = {6 * pidx(l, i) + pidx(k, j), 6 * pidx(i, k) + pidx(l, j),
                               6 * pidx(l, j) + pidx(i, k), 6 * pidx(j, k) + pidx(i, l)};
            for (int t = 0; t < 4; t++) {
                const int e = xy[t];
                int slot = 0;
                while (slot < 4 && m.c[e][slot] != 0 && !(m.P[e][slot] == P && m.R[e][slot] == R)) slot++;
                m.P[e][slot] = (signed char)P;
                m.R[e][slot] = (signed char)R;
                m.c[e][slot]++;
            }
        }
    return m;
}
constexpr H2Dst kH2 = make_h2dst();

struct ODist {
    double a2[6];   // a2 pairs (xx yy zz yz xz xy)
    double u[15];   // a4 index-sorted entries
};

// a2 and a4 body of the l <= 4 block x (x[0] = n00 > 0)
__host__ __device__ __forceinline__ void odist(const double (&x)[15], ODist& d) {
    const double r = 1.0 / x[0];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 15; q++)
            if (kL4Maps.a2[i][q] != 0.0) s += kL4Maps.a2[i][q] * x[q];
        d.a2[i] = s * r + (i < 3 ? 1.0 / 3.0 : 0.0);
    }
#pragma unroll
    for (int e = 0; e < 15; e++) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 15; q++)
            if (kL4Maps.a4[e][q] != 0.0) s += kL4Maps.a4[e][q] * x[q];
        d.u[e] = s * r;
    }
}

__host__ __device__ __forceinline__ double a2dot(const ODist& b, const ODist& n) {
    double s = 0.0;
#pragma unroll
    for (int P = 0; P < 6; P++) s += pw(P) * (b.a2[P] * n.a2[P]);
    return s;
}

// pair entry [X][Y] of a4_joint(b, n) times (1 - a2b : a2n)
__host__ __device__ __forceinline__ double joint_raw(const ODist& b, const ODist& n, int X, int Y) {
    double d = 0.0, o = 0.0;
#pragma unroll
    for (int R = 0; R < 3; R++) d += b.u[sxy(X, R)] * n.u[sxy(Y, R)];
#pragma unroll
    for (int R = 3; R < 6; R++) o += b.u[sxy(X, R)] * n.u[sxy(Y, R)];
    return (b.a2[X] * n.a2[Y] - d) - 2.0 * o;
}

// pair entry [X][Y] of a4_jointcross(b, n) = <n n v v> times (1 - a2b : a2n)
__host__ __device__ __forceinline__ double jointcross_raw(const ODist& b, const ODist& n, int X, int Y) {
    const int k = pa(Y), l = pb(Y);
    double s = 0.0;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++)
#pragma unroll
            for (int r = 0; r < 3; r++)
#pragma unroll
                for (int t = 0; t < 3; t++) {
                    const int e = levi(k, p, q) * levi(l, r, t);
                    if (e != 0)
                        s += (double)e * (n.u[kA4Slots.sym[27 * pa(X) + 9 * pb(X) + 3 * q + t]] * b.a2[pidx(p, r)]);
                }
    return s;
}

// <v v> of v = b x n (pair P) times (1 - a2b : a2n)
__host__ __device__ __forceinline__ double orth_a2_raw(const ODist& b, const ODist& n, int P) {
    const int i = pa(P), j = pb(P);
    double s = 0.0;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++)
#pragma unroll
            for (int r = 0; r < 3; r++)
#pragma unroll
                for (int t = 0; t < 3; t++) {
                    const int e = levi(i, p, q) * levi(j, r, t);
                    if (e != 0) s += (double)e * (b.a2[pidx(p, r)] * n.a2[pidx(q, t)]);
                }
    return s;
}

// indices (a0 <= a1 <= a2 <= a3) of sorted entry e and its multiplicity
// 4! / (n0! n1! n2!) among the 81 index tuples (tables, so that uses inside
// unrolled loops fold)
struct SortedIdx {
    int i[15][4];
    double mult[15];
};
__host__ __device__ constexpr SortedIdx make_sorted_idx() {
    SortedIdx t{};
    int c = 0;
    for (int a0 = 0; a0 < 3; a0++)
        for (int a1 = a0; a1 < 3; a1++)
            for (int a2 = a1; a2 < 3; a2++)
                for (int a3 = a2; a3 < 3; a3++) {
                    t.i[c][0] = a0, t.i[c][1] = a1, t.i[c][2] = a2, t.i[c][3] = a3;
                    int n[3] = {0, 0, 0};
                    n[a0]++, n[a1]++, n[a2]++, n[a3]++;
                    const int f[5] = {1, 1, 2, 6, 24};
                    t.mult[c] = 24.0 / (f[n[0]] * f[n[1]] * f[n[2]]);
                    c++;
                }
    return t;
}
constexpr SortedIdx kSorted = make_sorted_idx();
__host__ __device__ constexpr int sslot(int a, int b, int c, int d) { return kA4Slots.sym[27 * a + 9 * b + 3 * c + d]; }

// sorted entry e of <v v v v> times its normalisation (1 - 2 a2b:a2n + a4b::a4n)
__host__ __device__ __forceinline__ double orth_a4_raw(const ODist& b, const ODist& n, int e) {
    const int i0 = kSorted.i[e][0], i1 = kSorted.i[e][1], i2 = kSorted.i[e][2], i3 = kSorted.i[e][3];
    double s = 0.0;
#pragma unroll
    for (int c0 = 0; c0 < 2; c0++)
#pragma unroll
        for (int c1 = 0; c1 < 2; c1++)
#pragma unroll
            for (int c2 = 0; c2 < 2; c2++)
#pragma unroll
                for (int c3 = 0; c3 < 2; c3++) {
                    // eps_{i p q}: (p, q) = (i+1, i+2) (+) or (i+2, i+1) (-)
                    const int p0 = (i0 + 1 + c0) % 3, q0 = (i0 + 2 - c0) % 3;
                    const int p1 = (i1 + 1 + c1) % 3, q1 = (i1 + 2 - c1) % 3;
                    const int p2 = (i2 + 1 + c2) % 3, q2 = (i2 + 2 - c2) % 3;
                    const int p3 = (i3 + 1 + c3) % 3, q3 = (i3 + 2 - c3) % 3;
                    const double sg = ((c0 + c1 + c2 + c3) & 1) ? -1.0 : 1.0;
                    s += sg * (b.u[sslot(p0, p1, p2, p3)] * n.u[sslot(q0, q1, q2, q3)]);
                }
    return s;
}
// a4b :: a4n over all 81 index tuples
__host__ __device__ __forceinline__ double a4dot(const ODist& b, const ODist& n) {
    double s = 0.0;
#pragma unroll
    for (int e = 0; e < 15; e++) s += kSorted.mult[e] * (b.u[e] * n.u[e]);
    return s;
}

// Q6 += sum over pair entries j = J[X][Y] * js of the joint tensor J:
// -cm/4 (J + J^T) + cm3 sym4(J); jf(X, Y) returns the unscaled entry
template <class JF>
__host__ __device__ __forceinline__ void q6_add_joint(double (&Q)[36], double cm, double cm3, double js, JF jf) {
    const double a = -0.25 * cm * js, h = 0.25 * cm3 * js;
#pragma unroll
    for (int X = 0; X < 6; X++)
#pragma unroll
        for (int Y = 0; Y < 6; Y++) {
            const double j = jf(X, Y);
            Q[6 * X + Y] += a * j;
            Q[6 * Y + X] += a * j;
#pragma unroll
            for (int t = 0; t < 4; t++)
                if (kH2.c[6 * X + Y][t] != 0)
                    Q[6 * kH2.P[6 * X + Y][t] + kH2.R[6 * X + Y][t]] += ((double)kH2.c[6 * X + Y][t] * h) * j;
        }
}

// Q6 += w S-part of one distribution's a4
__host__ __device__ __forceinline__ void q6_add_a4(double (&Q)[36], double w, const double (&u)[15]) {
#pragma unroll
    for (int P = 0; P < 6; P++)
#pragma unroll
        for (int R = 0; R < 6; R++) Q[6 * P + R] += w * u[sxy(P, R)];
}

// body a4 excess at (3,2,1,2) (1-based) over the symmetric entry
__host__ __device__ __forceinline__ double quirk(const ODist& d) {
    return d.u[kA4Slots.s[27 * 2 + 9 * 1 + 0 + 1]] - d.u[kA4Slots.sym[27 * 2 + 9 * 1 + 0 + 1]];
}

struct JointF {
    const ODist& b;
    const ODist& n;
    __host__ __device__ __forceinline__ double operator()(int X, int Y) const { return joint_raw(b, n, X, Y); }
};
struct CrossF {
    const ODist& b;
    const ODist& n;
    __host__ __device__ __forceinline__ double operator()(int X, int Y) const { return jointcross_raw(b, n, X, Y); }
};

// Q6 and dq of the given branch (q1, q2, q3 all distributions)
__host__ __device__ __forceinline__ void ortho_q6_given(const ODist& d0, const ODist& d1, const ODist& d2,
                                                        const double* ci, double (&Q)[36], double& dq) {
    const double cs = ci[0] + ci[1] + ci[2];
    const double w0 = (cs - ci[0]) * 0.25, w1 = (cs - ci[1]) * 0.25, w2 = (cs - ci[2]) * 0.25;
#pragma unroll
    for (int t = 0; t < 36; t++) Q[t] = 0.0;
    q6_add_a4(Q, w0, d0.u);
    q6_add_a4(Q, w1, d1.u);
    q6_add_a4(Q, w2, d2.u);
    dq = w0 * quirk(d0) + w1 * quirk(d1) + w2 * quirk(d2);
    q6_add_joint(Q, ci[2], ci[5], 1.0 / (1.0 - a2dot(d0, d1)), JointF{d0, d1});   // <q1 q1 q2 q2>
    q6_add_joint(Q, ci[0], ci[3], 1.0 / (1.0 - a2dot(d1, d2)), JointF{d1, d2});   // <q2 q2 q3 q3>
    q6_add_joint(Q, ci[1], ci[4], 1.0 / (1.0 - a2dot(d0, d2)), JointF{d0, d2});   // <q1 q1 q3 q3>
}

// Q6 and dq of the derived branch (v = b x n)
__host__ __device__ __forceinline__ void ortho_q6_derived(const ODist& d0, const ODist& d1, const double* ci,
                                                          double (&Q)[36], double& dq) {
    const double cs = ci[0] + ci[1] + ci[2];
    const double w0 = (cs - ci[0]) * 0.25, w1 = (cs - ci[1]) * 0.25, w2 = (cs - ci[2]) * 0.25;
#pragma unroll
    for (int t = 0; t < 36; t++) Q[t] = 0.0;
    q6_add_a4(Q, w0, d0.u);
    q6_add_a4(Q, w1, d1.u);
    const double s01 = 1.0 / (1.0 - a2dot(d0, d1));
    const double n4 = 1.0 / (1.0 - 2.0 * a2dot(d0, d1) + a4dot(d0, d1));
    {
        double u2[15];
#pragma unroll
        for (int e = 0; e < 15; e++) u2[e] = orth_a4_raw(d0, d1, e) * n4;
        q6_add_a4(Q, w2, u2);
    }
    dq = w0 * quirk(d0) + w1 * quirk(d1);
    q6_add_joint(Q, ci[2], ci[5], s01, JointF{d0, d1});   // <q1 q1 q2 q2>
    q6_add_joint(Q, ci[0], ci[3], s01, CrossF{d1, d0});   // a4_jointcross(q2, q1)
    q6_add_joint(Q, ci[1], ci[4], s01, CrossF{d0, d1});   // a4_jointcross(q1, q2)
}

// eps(tau) : vw = sum Q_lkij tau_ji (vw)_kl of one stress state
__host__ __device__ __forceinline__ double q6_state(const double (&Q)[36], double dq, const double (&tau)[9],
                                                    const double* v, const double* w) {
    double t6[6];
#pragma unroll
    for (int R = 0; R < 6; R++) t6[R] = pw(R) * tau[3 * pa(R) + pb(R)];
    double r = dq * tau[3] * (v[1] * w[2]);
#pragma unroll
    for (int P = 0; P < 6; P++) {
        double y = 0.0;
#pragma unroll
        for (int R = 0; R < 6; R++) y += Q[6 * P + R] * t6[R];
        const int a = pa(P), b = pb(P);
        r += (P < 3 ? v[a] * w[a] : v[a] * w[b] + v[b] * w[a]) * y;
    }
    return r;
}

// the six stress states of the frame e (rows e1, e2, e3): tau_vv(e_i) =
// I/3 - e_i e_i, tau_vw = v w + w v (src/enhancementfactors.f90:398-413)
__host__ __device__ __forceinline__ void ortho_eij_frame(const double (&Q)[36], double dq, const double (&Qi)[36],
                                                         double dqi, const double (&e)[9], double* out) {
#pragma unroll 1
    for (int s = 0; s < 6; s++) {
        const int iv = s < 3 ? s : (s == 3 ? 1 : 0), iw = s < 3 ? s : (s == 5 ? 1 : 2);
        double v[3], w[3], tau[9];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            v[i] = iv == 0 ? e[i] : (iv == 1 ? e[3 + i] : e[6 + i]);
            w[i] = iw == 0 ? e[i] : (iw == 1 ? e[3 + i] : e[6 + i]);
        }
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                tau[3 * i + j] = s < 3 ? (i == j ? 1.0 / 3.0 : 0.0) - v[i] * v[j] : v[i] * w[j] + w[i] * v[j];
        out[s] = q6_state(Q, dq, tau, v, w) / q6_state(Qi, dqi, tau, v, w);
    }
}

// Cartesian frame: E_ii = Q6 row i against (I/3 - e_i e_i), E_vw = Q6[P][P]
// (the dq entry meets no nonzero tau there)
__host__ __device__ __forceinline__ double q6_cart(const double (&Q)[36], int s) {
    if (s >= 3) return Q[7 * s];
    return (Q[6 * s + 0] + Q[6 * s + 1] + Q[6 * s + 2]) * (1.0 / 3.0) - Q[7 * s];
}

__device__ __forceinline__ void load_l4(const double* __restrict__ x, int64_t P, int64_t p, int nq, double (&xr)[15]) {
#pragma unroll
    for (int r = 0; r < 15; r++) xr[r] = r < nq ? x[(int64_t)r * P + p] : 0.0;
}

}  // namespace

void ortho_q6_iso(const double ci[6], double Q6[36], double* dq) {
    double x[15] = {1.0};
    ODist d;
    odist(x, d);
    double Q[36];
    ortho_q6_given(d, d, d, ci, Q, *dq);
    for (int t = 0; t < 36; t++) Q6[t] = Q[t];
}

// Q6 and dq of parcel p: MODE 0: v given when its n00 is above 1e-8
// (src/moments.f90:371-383), tested per parcel; MODE 1: v = b x n for all
template <int MODE>
__device__ __forceinline__ void ortho_q6_parcel(const OrthoArgs& a, int64_t p, double (&Q)[36], double& dq) {
    // all three blocks requested at once (one memory round trip); q3's n00
    // decides the branch
    double x3r[15];
    if (MODE == 0) load_l4(a.x3, a.P, p, a.nq, x3r);
    const bool given = MODE == 0 && x3r[0] > 1e-8;
    if (given) {
        // ortho_q6_given in the order q1 q2 | q3 (q2 dropped) | q1 again, so
        // that at most two distributions are live at a time
        const double* ci = a.ci;
        const double cs = ci[0] + ci[1] + ci[2];
        const double w0 = (cs - ci[0]) * 0.25, w1 = (cs - ci[1]) * 0.25, w2 = (cs - ci[2]) * 0.25;
#pragma unroll
        for (int t = 0; t < 36; t++) Q[t] = 0.0;
        double xr[15];
        ODist d0, d2;
        // q1's block kept in LDS for its second evaluation (one 15-row column
        // per lane: read back only by the lane that wrote it, so no barrier)
        __shared__ double q1s[15][256];
        {
            load_l4(a.x1, a.P, p, a.nq, xr);
#pragma unroll
            for (int r = 0; r < 15; r++) q1s[r][threadIdx.x] = xr[r];
            odist(xr, d0);
            ODist d1;
            load_l4(a.x2, a.P, p, a.nq, xr);
            odist(xr, d1);
            q6_add_a4(Q, w0, d0.u);
            q6_add_a4(Q, w1, d1.u);
            dq = w0 * quirk(d0) + w1 * quirk(d1);
            q6_add_joint(Q, ci[2], ci[5], 1.0 / (1.0 - a2dot(d0, d1)), JointF{d0, d1});
            odist(x3r, d2);
            q6_add_a4(Q, w2, d2.u);
            dq += w2 * quirk(d2);
            q6_add_joint(Q, ci[0], ci[3], 1.0 / (1.0 - a2dot(d1, d2)), JointF{d1, d2});
        }
        // q1 evaluated again from LDS; the compiler barrier keeps it from
        // forwarding the stored registers, so the first evaluation need not
        // stay live
        asm volatile("" ::: "memory");
#pragma unroll
        for (int r = 0; r < 15; r++) xr[r] = q1s[r][threadIdx.x];
        odist(xr, d0);
        q6_add_joint(Q, ci[1], ci[4], 1.0 / (1.0 - a2dot(d0, d2)), JointF{d0, d2});
    } else {
        double xr[15];
        load_l4(a.x1, a.P, p, a.nq, xr);
        ODist d0;
        odist(xr, d0);
        load_l4(a.x2, a.P, p, a.nq, xr);
        ODist d1;
        odist(xr, d1);
        ortho_q6_derived(d0, d1, a.ci, Q, dq);
    }
}

// Eij of orthotropic grains (Sachs; src/enhancementfactors.f90:134-188): one
// lane per parcel, MODE as for ortho_q6_parcel
template <bool CART, int MODE>
__global__ void __launch_bounds__(256) eij_ortho_kernel(OrthoArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.P) return;
    double Q[36], dq;
    ortho_q6_parcel<MODE>(a, p, Q, dq);
    double* o = a.Eij + p * 6;
    if (CART) {
#pragma unroll
        for (int s = 0; s < 6; s++) o[s] = q6_cart(Q, s) / q6_cart(a.q6iso, s);
    } else {
        double e[9];
#pragma unroll
        for (int i = 0; i < 9; i++) e[i] = a.frames[p * 9 + i];
        ortho_eij_frame(Q, dq, a.q6iso, a.dqiso, e, o);
    }
}

// Evw of orthotropic grains for the given stress states (v, w, tau), shared
// by all parcels (src/enhancementfactors.f90:134-160, the reference's
// Evw_orthotropic): one lane per parcel
template <int MODE>
__global__ void __launch_bounds__(256) evw_ortho_kernel(OrthoArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.P) return;
    double Q[36], dq;
    ortho_q6_parcel<MODE>(a, p, Q, dq);
    double* o = a.Eij + p * a.nstates;
#pragma unroll 1
    for (int q = 0; q < a.nstates; q++) {
        const double* st = a.states + 15 * q;
        double tau[9];
#pragma unroll
        for (int i = 0; i < 9; i++) tau[i] = st[6 + i];
        o[q] = q6_state(Q, dq, tau, st, st + 3) / q6_state(a.q6iso, a.dqiso, tau, st, st + 3);
    }
}

// the l <= 4 blocks of NS states rotated by one direction's real blocks R
// (R2 5x5 then R4 9x9, row-major; the same R for the whole workgroup, so its
// entries are scalar loads, each used for all NS states at once)
template <int NS>
__device__ __forceinline__ void rotate_l4(const double* __restrict__ R, double (&x)[3][15]) {
    double y[NS][15];
#pragma unroll
    for (int s = 0; s < NS; s++) y[s][0] = x[s][0];
#pragma unroll
    for (int i = 0; i < 5; i++)
#pragma unroll
        for (int s = 0; s < NS; s++) {
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < 5; j++) acc += R[5 * i + j] * x[s][1 + j];
            y[s][1 + i] = acc;
        }
#pragma unroll
    for (int i = 0; i < 9; i++)
#pragma unroll
        for (int s = 0; s < NS; s++) {
            double acc = 0.0;
#pragma unroll
            for (int j = 0; j < 9; j++) acc += R[25 + 9 * i + j] * x[s][6 + j];
            y[s][6 + i] = acc;
        }
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
        for (int q = 0; q < 15; q++) x[s][q] = y[s][q];
}

// Qn += mu/2 (a2 + a2.kk + kk.a2 + a2_zz I) + lam a4(i,j,z,z) of one distribution
__device__ __forceinline__ void qn_add_dist(double (&Qn)[9], const double (&a2)[6], const double (&zz)[6], double mu,
                                            double lam) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const double anti = (j == 2 ? a2[pidx(i, 2)] : 0.0) + (i == 2 ? a2[pidx(2, j)] : 0.0);
            Qn[3 * i + j] += mu * 0.5 * (a2[pidx(i, j)] + anti + (i == j ? a2[2] : 0.0)) + lam * zz[pidx(i, j)];
        }
}

// Qn += lam' (J(z,i,j,z) + J(j,z,z,i)) of one joint tensor (pair form, scaled by js)
template <class JF>
__device__ __forceinline__ void qn_add_joint(double (&Qn)[9], double lam3, double js, JF jf) {
    const double c = lam3 * js;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) Qn[3 * i + j] += c * (jf(pidx(2, i), pidx(j, 2)) + jf(pidx(j, 2), pidx(2, i)));
}

// a4(i, j, z, z) pairs of a distribution's body a4
__device__ __forceinline__ void a4_zz(const ODist& d, double (&zz)[6]) {
#pragma unroll
    for (int P = 0; P < 6; P++) zz[P] = d.u[kA4Slots.s[27 * pa(P) + 9 * pb(P) + 3 * 2 + 2]];
}

// Workgroup w of a one-dimensional grid of n is dispatched to XCD w % 8
// (round robin over the eight XCDs); its index in an order that gives each
// XCD one contiguous run of [0, n), in dispatch order.  A bijection for any
// n, so the mapping is correct whatever the dispatcher does; it only decides
// which workgroups share an XCD's L2.
__device__ __forceinline__ int64_t xcd_contiguous(int64_t w, int64_t n) {
    const int64_t q = n >> 3, r = n & 7, x = w & 7;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + (w >> 3);
}

// elastic Vi of orthotropic grains (Voigt; src/wavepropagation.f90:41-68,
// 98-115): one lane per parcel, one direction per workgroup; the
// distributions rotated as rotate_nlm4 rotates them, the acoustic tensor for
// k = z
//   Qn = sum_m mu_m/2 (a2 + a2.kk + kk.a2 + (a2:kk) I) + lam_m a4_m : kk
//        + lam_m+3 (2 sym2(J_m)) first-last : kk
// and the reference's Cardano roots.  The grid is (parcel block, direction)
// with the direction fastest, laid out so that the ndir workgroups of one
// parcel block run on one XCD one after another: every direction after the
// first reads the parcels' blocks from that XCD's L2.  MODE as for
// eij_ortho_kernel.
template <int MODE>
__global__ void __launch_bounds__(256) vi_ortho_kernel(OrthoArgs a) {
    const int64_t lw = xcd_contiguous(blockIdx.x, gridDim.x);
    const int64_t blk = lw / a.ndir;
    const int d = (int)(lw - blk * a.ndir);
    const int64_t p = blk * blockDim.x + threadIdx.x;
    if (p >= a.P) return;
    const double* __restrict__ R = a.R + (int64_t)d * ROT4_STRIDE;
    const bool given = MODE == 0 && a.x3[p] > 1e-8;   // n00 is invariant under rotation
    double x[3][15];
    load_l4(a.x1, a.P, p, a.nq, x[0]);
    load_l4(a.x2, a.P, p, a.nq, x[1]);
    double Qn[9];
#pragma unroll
    for (int i = 0; i < 9; i++) Qn[i] = 0.0;
    if (given) {
        load_l4(a.x3, a.P, p, a.nq, x[2]);
        rotate_l4<3>(R, x);
        ODist q0, q1, q2;
        odist(x[0], q0);
        odist(x[1], q1);
        odist(x[2], q2);
        double zz[6];
        a4_zz(q0, zz);
        qn_add_dist(Qn, q0.a2, zz, a.mu[0], a.lam[0]);
        a4_zz(q1, zz);
        qn_add_dist(Qn, q1.a2, zz, a.mu[1], a.lam[1]);
        a4_zz(q2, zz);
        qn_add_dist(Qn, q2.a2, zz, a.mu[2], a.lam[2]);
        qn_add_joint(Qn, a.lam[5], 1.0 / (1.0 - a2dot(q0, q1)), JointF{q0, q1});   // <q1 q1 q2 q2>
        qn_add_joint(Qn, a.lam[3], 1.0 / (1.0 - a2dot(q1, q2)), JointF{q1, q2});   // <q2 q2 q3 q3>
        qn_add_joint(Qn, a.lam[4], 1.0 / (1.0 - a2dot(q0, q2)), JointF{q0, q2});   // <q1 q1 q3 q3>
    } else {
        rotate_l4<2>(R, x);
        ODist q0, q1;
        odist(x[0], q0);
        odist(x[1], q1);
        double zz[6];
        a4_zz(q0, zz);
        qn_add_dist(Qn, q0.a2, zz, a.mu[0], a.lam[0]);
        a4_zz(q1, zz);
        qn_add_dist(Qn, q1.a2, zz, a.mu[1], a.lam[1]);
        const double s01 = 1.0 / (1.0 - a2dot(q0, q1));
        const double n4 = 1.0 / (1.0 - 2.0 * a2dot(q0, q1) + a4dot(q0, q1));
        double a2v[6];
#pragma unroll
        for (int P = 0; P < 6; P++) a2v[P] = orth_a2_raw(q0, q1, P) * s01;
#pragma unroll
        for (int P = 0; P < 6; P++) zz[P] = orth_a4_raw(q0, q1, kA4Slots.sym[27 * pa(P) + 9 * pb(P) + 3 * 2 + 2]) * n4;
        qn_add_dist(Qn, a2v, zz, a.mu[2], a.lam[2]);
        qn_add_joint(Qn, a.lam[5], s01, JointF{q0, q1});   // <q1 q1 q2 q2>
        qn_add_joint(Qn, a.lam[3], s01, CrossF{q1, q0});   // a4_jointcross(q2, q1)
        qn_add_joint(Qn, a.lam[4], s01, CrossF{q0, q1});   // a4_jointcross(q1, q2)
    }
    double vi[3];
    qnorm_to_vi(Qn, a.rho, vi);
    double* o = a.Vi + (p * a.ndir + d) * 3;
    o[0] = vi[0];
    o[1] = vi[1];
    o[2] = vi[2];
}

// ---- discrete grains (src/moments.f90:386-417, src/enhancementfactors.f90:191-266,
// src/wavepropagation.f90:70-96) ----

#ifndef SFG_HOST_EMUL   // wave shuffles: not emulated lane by lane
// grain sums of one ensemble chunk: blockIdx.x = e * G + chunk, blockIdx.y = m;
// per grain the axis m's pair products (a2), their products (a4) and the
// pair products of the two other axes (J_m); wave shuffles, then the four
// waves' sums through LDS
__global__ void __launch_bounds__(256) ortho_moments_kernel(OrthoDiscArgs a) {
    __shared__ double red[4][ORTHO_NMOM];
    const int m = (int)blockIdx.y;
    const int64_t e = blockIdx.x / a.G;
    const int g = (int)(blockIdx.x % a.G);
    const int64_t per = (a.N + a.G - 1) / a.G;
    const int64_t g0 = g * per, g1 = g0 + per < a.N ? g0 + per : a.N;
    const double* ax[3] = {a.b, a.n, a.v};
    const double* __restrict__ qm = ax[m] + e * a.N * 3;
    const double* __restrict__ qa = ax[m == 0 ? 1 : 0] + e * a.N * 3;
    const double* __restrict__ qb = ax[m == 2 ? 1 : 2] + e * a.N * 3;
    double acc[ORTHO_NMOM];
#pragma unroll
    for (int t = 0; t < ORTHO_NMOM; t++) acc[t] = 0.0;
    for (int64_t i = g0 + threadIdx.x; i < g1; i += 256) {
        double q[3], x[3], y[3];
#pragma unroll
        for (int c = 0; c < 3; c++) q[c] = qm[i * 3 + c], x[c] = qa[i * 3 + c], y[c] = qb[i * 3 + c];
        double p2[6], px[6], py[6];
#pragma unroll
        for (int P = 0; P < 6; P++) {
            p2[P] = q[pa(P)] * q[pb(P)];
            px[P] = x[pa(P)] * x[pb(P)];
            py[P] = y[pa(P)] * y[pb(P)];
        }
#pragma unroll
        for (int P = 0; P < 6; P++) acc[P] += p2[P];
#pragma unroll
        for (int t = 0; t < 15; t++)
            acc[6 + t] += p2[pidx(kSorted.i[t][0], kSorted.i[t][1])] * p2[pidx(kSorted.i[t][2], kSorted.i[t][3])];
#pragma unroll
        for (int X = 0; X < 6; X++)
#pragma unroll
            for (int Y = 0; Y < 6; Y++) acc[21 + 6 * X + Y] += px[X] * py[Y];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int t = 0; t < ORTHO_NMOM; t++) {
        double v = acc[t];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) red[w][t] = v;
    }
    __syncthreads();
    if (threadIdx.x < ORTHO_NMOM) {
        const int t = threadIdx.x;
        a.part[((e * a.G + g) * 3 + m) * ORTHO_NMOM + t] = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
    }
}

// the G partial sums of each (ensemble, m, moment) in chunk order, over N
__global__ void __launch_bounds__(256) ortho_moments_final_kernel(OrthoDiscArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.E * 3 * ORTHO_NMOM) return;
    const int64_t e = t / (3 * ORTHO_NMOM);
    const int r = (int)(t % (3 * ORTHO_NMOM));
    double s = 0.0;
    for (int g = 0; g < a.G; g++) s += a.part[(e * a.G + g) * 3 * ORTHO_NMOM + r];
    a.mom[t] = s / (double)a.N;
}
#endif

// Eij of an ensemble's grain averages: one lane per ensemble
__global__ void __launch_bounds__(64) eij_ortho_mom_kernel(OrthoMomArgs a) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.E) return;
    const double* M = a.mom + e * 3 * ORTHO_NMOM;
    const double cs = a.ci[0] + a.ci[1] + a.ci[2];
    double Q[36];
#pragma unroll
    for (int t = 0; t < 36; t++) Q[t] = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double u[15];
#pragma unroll
        for (int t = 0; t < 15; t++) u[t] = M[d * ORTHO_NMOM + 6 + t];
        q6_add_a4(Q, (cs - a.ci[d]) * 0.25, u);
    }
#pragma unroll
    for (int m = 0; m < 3; m++) {
        const double* J = M + m * ORTHO_NMOM + 21;
        q6_add_joint(Q, a.ci[m], a.ci[m + 3], 1.0, [&](int X, int Y) { return J[6 * X + Y]; });
    }
    double* o = a.Eij + e * 6;
    if (!a.frames) {
#pragma unroll
        for (int s = 0; s < 6; s++) o[s] = q6_cart(Q, s) / q6_cart(a.q6iso, s);
    } else {
        double f[9];
#pragma unroll
        for (int i = 0; i < 9; i++) f[i] = a.frames[e * 9 + i];
        ortho_eij_frame(Q, 0.0, a.q6iso, a.dqiso, f, o);
    }
}

// Vi of an ensemble's grain averages along kdir[blockIdx.y]: the reference
// rotates the grains so that the direction is z and forms the acoustic
// tensor for k = z; here the tensor is formed in the grains' frame for k =
// kdir, the same tensor in the other basis (the rotation is orthogonal), whose
// invariants -- all that the Cardano roots read -- are the same
__global__ void __launch_bounds__(64) vi_ortho_mom_kernel(OrthoMomArgs a) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.E) return;
    const int d = (int)blockIdx.y;
    const double k[3] = {a.kdir[3 * d], a.kdir[3 * d + 1], a.kdir[3 * d + 2]};
    const double* M = a.mom + e * 3 * ORTHO_NMOM;
    double Qn[9];
#pragma unroll
    for (int i = 0; i < 9; i++) Qn[i] = 0.0;
#pragma unroll 1
    for (int m = 0; m < 3; m++) {
        const double* a2 = M + m * ORTHO_NMOM;
        const double* u = a2 + 6;
        const double* J = a2 + 21;
        double a2k[3];
#pragma unroll
        for (int i = 0; i < 3; i++) a2k[i] = a2[pidx(i, 0)] * k[0] + a2[pidx(i, 1)] * k[1] + a2[pidx(i, 2)] * k[2];
        const double kak = a2k[0] * k[0] + a2k[1] * k[1] + a2k[2] * k[2];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                // mu/2 (a2 + a2.kk + kk.a2 + (k.a2.k) I) + lam a4 : kk
                double q4 = 0.0, jf = 0.0;
#pragma unroll
                for (int p = 0; p < 3; p++)
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        const double kk = k[p] * k[q];
                        q4 += u[kA4Slots.sym[27 * i + 9 * j + 3 * p + q]] * kk;
                        // lam' (J(p,i,j,q) + J(j,q,p,i)) k_p k_q
                        jf += (J[6 * pidx(p, i) + pidx(j, q)] + J[6 * pidx(j, q) + pidx(p, i)]) * kk;
                    }
                Qn[3 * i + j] += a.mu[m] * 0.5 * (a2[pidx(i, j)] + a2k[i] * k[j] + k[i] * a2k[j] + (i == j ? kak : 0.0)) +
                                 a.lam[m] * q4 + a.lam[m + 3] * jf;
            }
    }
    double vi[3];
    qnorm_to_vi(Qn, a.rho, vi);
    double* o = a.Vi + (e * a.ndir + d) * 3;
    o[0] = vi[0];
    o[1] = vi[1];
    o[2] = vi[2];
}

#ifdef SFG_HOST_EMUL
template void eij_ortho_kernel<false, 0>(OrthoArgs);
template void eij_ortho_kernel<false, 1>(OrthoArgs);
template void eij_ortho_kernel<true, 0>(OrthoArgs);
template void eij_ortho_kernel<true, 1>(OrthoArgs);
template void vi_ortho_kernel<0>(OrthoArgs);
template void vi_ortho_kernel<1>(OrthoArgs);
template void evw_ortho_kernel<0>(OrthoArgs);
template void evw_ortho_kernel<1>(OrthoArgs);
#endif

#ifndef SFG_HOST_EMUL
static unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_vi_elastic(const ViElasticArgs& a, hipStream_t s) {
    const int64_t n = (a.idx ? a.nidx : a.P) * a.ndir;
    if (n == 0) return hipSuccess;
    if (a.idx && a.nidx_dev)   // a device-side count: at most 2048 workgroups
        hipLaunchKernelGGL(vi_elastic_list_kernel, dim3(std::min(nblk(n, 128), 2048u)), dim3(128), 0, s, a);
    else
        hipLaunchKernelGGL(vi_elastic_kernel, dim3(nblk(n, 128)), dim3(128), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_vi_em(const ViEMArgs& a, hipStream_t s) {
    if (a.P * a.ndir == 0) return hipSuccess;
    hipLaunchKernelGGL(vi_em_kernel, dim3(nblk(a.P * a.ndir, 128)), dim3(128), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_eij_ortho(const OrthoArgs& args, hipStream_t s) {
    if (args.P == 0) return hipSuccess;
    OrthoArgs a = args;
    ortho_q6_iso(a.ci, a.q6iso, &a.dqiso);
    const dim3 g(nblk(a.P, 256)), b(256);
    if (a.frames) {
        if (a.x3) hipLaunchKernelGGL((eij_ortho_kernel<false, 0>), g, b, 0, s, a);
        else hipLaunchKernelGGL((eij_ortho_kernel<false, 1>), g, b, 0, s, a);
    } else {
        if (a.x3) hipLaunchKernelGGL((eij_ortho_kernel<true, 0>), g, b, 0, s, a);
        else hipLaunchKernelGGL((eij_ortho_kernel<true, 1>), g, b, 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_evw_ortho(const OrthoArgs& args, hipStream_t s) {
    if (args.P == 0 || args.nstates == 0) return hipSuccess;
    OrthoArgs a = args;
    ortho_q6_iso(a.ci, a.q6iso, &a.dqiso);
    const dim3 g(nblk(a.P, 256)), b(256);
    if (a.x3) hipLaunchKernelGGL((evw_ortho_kernel<0>), g, b, 0, s, a);
    else hipLaunchKernelGGL((evw_ortho_kernel<1>), g, b, 0, s, a);
    return hipGetLastError();
}

hipError_t launch_vi_ortho(const OrthoArgs& a, hipStream_t s) {
    if (a.P * a.ndir == 0) return hipSuccess;
    const int64_t nw = (int64_t)nblk(a.P, 256) * a.ndir;   // (parcel block, direction) workgroups
    if (nw > 0x7fffffff) return hipErrorInvalidValue;
    const dim3 g((unsigned)nw), b(256);
    if (a.x3) hipLaunchKernelGGL((vi_ortho_kernel<0>), g, b, 0, s, a);
    else hipLaunchKernelGGL((vi_ortho_kernel<1>), g, b, 0, s, a);
    return hipGetLastError();
}

hipError_t launch_ortho_moments(const OrthoDiscArgs& a, hipStream_t s) {
    if (a.E == 0) return hipSuccess;
    hipLaunchKernelGGL(ortho_moments_kernel, dim3((unsigned)(a.E * a.G), 3), dim3(256), 0, s, a);
    hipLaunchKernelGGL(ortho_moments_final_kernel, dim3(nblk(a.E * 3 * ORTHO_NMOM, 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_eij_ortho_mom(const OrthoMomArgs& args, hipStream_t s) {
    if (args.E == 0) return hipSuccess;
    OrthoMomArgs a = args;
    ortho_q6_iso(a.ci, a.q6iso, &a.dqiso);
    hipLaunchKernelGGL(eij_ortho_mom_kernel, dim3(nblk(a.E, 64)), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_vi_ortho_mom(const OrthoMomArgs& a, hipStream_t s) {
    if (a.E == 0 || a.ndir == 0) return hipSuccess;
    if (a.ndir > 65535) return hipErrorInvalidValue;   // one grid row per direction
    hipLaunchKernelGGL(vi_ortho_mom_kernel, dim3(nblk(a.E, 64), (unsigned)a.ndir), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_a4full(const A4FullArgs& a, hipStream_t s) {
    if (a.P == 0) return hipSuccess;
    hipLaunchKernelGGL(a4full_kernel, dim3(nblk(a.P, 128)), dim3(128), 0, s, a);
    return hipGetLastError();
}
#endif

}  // namespace sfg
