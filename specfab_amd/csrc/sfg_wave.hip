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

#include "sfg_a4slots.h"
#include "sfg_device.h"
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
// (the reference's a2 body, through the Mandel map) and its full a4 (the
// REAL(4) body, its (3,2,1,2) entry included); the joint tensors follow from
// the two distributions' moments under the product density with the kernel
// sin^2 of the axes' angle, as the reference's generated bodies evaluate them
// (src/moments.f90:242-310; identified against them to their REAL(4) level):
//   a4_joint(b, n)      = (a2b (x) a2n - a4b : a4n) / (1 - a2b : a2n)       <b b n n>
//   a4_jointcross(b, n) = eps eps (a4n (x) a2b) / (1 - a2b : a2n)           <n n v v>
//   a2_orth(b, n)       = eps eps (a2b (x) a2n) / (1 - a2b : a2n)           <v v>
//   a4_orth(b, n)       = eps^4 (a4b (x) a4n) / <sin^4> = ... / (1 - 2 a2b:a2n + a4b::a4n)
// with v = b x n and eps the Levi-Civita symbol.

namespace {

struct OrthoDist {
    double a2[9];    // a2 (matrix)
    double u[15];    // a4 body, index-sorted entries
    bool sym;        // a4 from a formula: symmetric (no (3,2,1,2) entry)
};

// pair index of (a, b), a symmetric pair: xx yy zz yz xz xy
__device__ __forceinline__ constexpr int pidx(int a, int b) { return a == b ? a : 6 - a - b; }

__device__ __forceinline__ double levi(int i, int j, int k) {
    return (i == j || j == k || i == k) ? 0.0 : (((j - i + 3) % 3 == 1) ? 1.0 : -1.0);
}

// a2 matrix and a4 body entries of the l <= 4 block xr (x0 = xr[0])
__device__ __forceinline__ void ortho_dist(const double* A2, const double* A4F, const double* xr, OrthoDist& d) {
    const double x0 = xr[0], r2i = 0.70710678118654752440;
    double a2v[6];
#pragma unroll
    for (int i = 0; i < 6; i++) a2v[i] = map_row(A2, i, xr, 15) / x0 + (i < 3 ? 1.0 / 3.0 : 0.0);
    d.a2[0] = a2v[0], d.a2[4] = a2v[1], d.a2[8] = a2v[2];
    d.a2[5] = d.a2[7] = a2v[3] * r2i, d.a2[2] = d.a2[6] = a2v[4] * r2i, d.a2[1] = d.a2[3] = a2v[5] * r2i;
#pragma unroll
    for (int e = 0; e < 15; e++) d.u[e] = map_row(A4F, e, xr, 15) / x0;
    d.sym = false;
}

__device__ __forceinline__ double a4e(const OrthoDist& d, int i, int j, int k, int l) {
    const int t = 27 * i + 9 * j + 3 * k + l;
    return d.sym ? d.u[kA4Slots.sym[t]] : d.u[kA4Slots.s[t]];
}
__device__ __forceinline__ double a4sym(const OrthoDist& d, int i, int j, int k, int l) {
    return d.u[kA4Slots.sym[27 * i + 9 * j + 3 * k + l]];
}
__device__ __forceinline__ double dd22(const double* A, const double* B) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 9; i++) s += A[i] * B[i];
    return s;
}

// J[P][R] (pair indices) of a4_joint(b, n)
__device__ __forceinline__ void a4_joint6(const OrthoDist& b, const OrthoDist& n, double (&J)[6][6]) {
    const double nrm = 1.0 - dd22(b.a2, n.a2);
#pragma unroll
    for (int P = 0; P < 6; P++)
#pragma unroll
        for (int R = 0; R < 6; R++) {
            const int i = P < 3 ? P : (P == 3 ? 1 : 0), j = P < 3 ? P : (P == 5 ? 1 : 2);
            const int k = R < 3 ? R : (R == 3 ? 1 : 0), l = R < 3 ? R : (R == 5 ? 1 : 2);
            double s = b.a2[3 * i + j] * n.a2[3 * k + l];
#pragma unroll
            for (int p = 0; p < 3; p++)
#pragma unroll
                for (int q = 0; q < 3; q++) s -= a4sym(b, i, j, p, q) * a4sym(n, k, l, p, q);
            J[P][R] = s / nrm;
        }
}

// J[P][R] of a4_jointcross(b, n) = <n n v v>
__device__ __forceinline__ void a4_jointcross6(const OrthoDist& b, const OrthoDist& n, double (&J)[6][6]) {
    const double nrm = 1.0 - dd22(b.a2, n.a2);
#pragma unroll
    for (int P = 0; P < 6; P++)
#pragma unroll
        for (int R = 0; R < 6; R++) {
            const int i = P < 3 ? P : (P == 3 ? 1 : 0), j = P < 3 ? P : (P == 5 ? 1 : 2);
            const int k = R < 3 ? R : (R == 3 ? 1 : 0), l = R < 3 ? R : (R == 5 ? 1 : 2);
            double s = 0.0;
#pragma unroll
            for (int p = 0; p < 3; p++)
#pragma unroll
                for (int q = 0; q < 3; q++)
#pragma unroll
                    for (int r = 0; r < 3; r++)
#pragma unroll
                        for (int t = 0; t < 3; t++) {
                            const double e = levi(k, p, q) * levi(l, r, t);
                            if (e != 0.0) s += e * a4sym(n, i, j, q, t) * b.a2[3 * p + r];
                        }
            J[P][R] = s / nrm;
        }
}

// <v v> and <v^4> of v = b x n: the derived third distribution
__device__ __forceinline__ void ortho_third(const OrthoDist& b, const OrthoDist& n, OrthoDist& v) {
    const double n2 = 1.0 - dd22(b.a2, n.a2);
    double a44 = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = 0; k < 3; k++)
#pragma unroll
                for (int l = 0; l < 3; l++) a44 += a4sym(b, i, j, k, l) * a4sym(n, i, j, k, l);
    const double n4 = 1.0 - 2.0 * dd22(b.a2, n.a2) + a44;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
#pragma unroll
            for (int p = 0; p < 3; p++)
#pragma unroll
                for (int q = 0; q < 3; q++)
#pragma unroll
                    for (int r = 0; r < 3; r++)
#pragma unroll
                        for (int t = 0; t < 3; t++) {
                            const double e = levi(i, p, q) * levi(j, r, t);
                            if (e != 0.0) s += e * b.a2[3 * p + r] * n.a2[3 * q + t];
                        }
            v.a2[3 * i + j] = s / n2;
        }
    // the 15 sorted entries (a0 <= a1 <= a2 <= a3) of <v^4>
    int e = 0;
#pragma unroll
    for (int a0 = 0; a0 < 3; a0++)
#pragma unroll
        for (int a1 = a0; a1 < 3; a1++)
#pragma unroll
            for (int a2 = a1; a2 < 3; a2++)
#pragma unroll
                for (int a3 = a2; a3 < 3; a3++) {
                    double s = 0.0;
#pragma unroll
                    for (int p0 = 0; p0 < 3; p0++)
#pragma unroll
                        for (int q0 = 0; q0 < 3; q0++) {
                            const double e0 = levi(a0, p0, q0);
                            if (e0 == 0.0) continue;
#pragma unroll
                            for (int p1 = 0; p1 < 3; p1++)
#pragma unroll
                                for (int q1 = 0; q1 < 3; q1++) {
                                    const double e1 = e0 * levi(a1, p1, q1);
                                    if (e1 == 0.0) continue;
#pragma unroll
                                    for (int p2 = 0; p2 < 3; p2++)
#pragma unroll
                                        for (int q2 = 0; q2 < 3; q2++) {
                                            const double e2 = e1 * levi(a2, p2, q2);
                                            if (e2 == 0.0) continue;
#pragma unroll
                                            for (int p3 = 0; p3 < 3; p3++)
#pragma unroll
                                                for (int q3 = 0; q3 < 3; q3++) {
                                                    const double e3 = e2 * levi(a3, p3, q3);
                                                    if (e3 == 0.0) continue;
                                                    s += e3 * a4sym(b, p0, p1, p2, p3) * a4sym(n, q0, q1, q2, q3);
                                                }
                                        }
                                }
                        }
                    v.u[e++] = s / n4;
                }
    v.sym = true;
}

// the a4_jk triple of ai_orthotropic (src/moments.f90:364-383): jk[0] = <q2 q2 q3 q3>
// (given) or a4_jointcross(q2, q1) (derived), jk[1] = <q1 q1 q3 q3> or
// a4_jointcross(q1, q2), jk[2] = <q1 q1 q2 q2>
__device__ __forceinline__ void ortho_joints(const OrthoDist* q, bool derived, int m, double (&J)[6][6]) {
    if (m == 2) a4_joint6(q[0], q[1], J);
    else if (!derived) a4_joint6(m == 0 ? q[1] : q[0], q[2], J);
    else if (m == 0) a4_jointcross6(q[1], q[0], J);
    else a4_jointcross6(q[0], q[1], J);
}

__device__ __forceinline__ double jt(const double (&J)[6][6], int a, int b, int c, int d) {
    return J[pidx(a, b)][pidx(c, d)];
}

// Q = sum_m ci_m M2_m + ci_{m+3} H2_m (src/homogenizations.f90:303-314), accumulated into Q
__device__ __forceinline__ void ortho_Q_add(const OrthoDist* q, int m, const double (&J)[6][6], const double* ci,
                                            double* Q, int stride) {
    const int jm = (m + 1) % 3, km = (m + 2) % 3;   // ji, ki of the reference (1-based [2,3,1], [3,1,2])
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const double M2 = (a4e(q[jm], l, k, i, j) + a4e(q[km], l, k, i, j) -
                                       2 * ((jt(J, l, k, i, j) + jt(J, i, j, l, k)) / 2)) / 4;
                    const double H2 = (jt(J, l, i, k, j) + jt(J, i, k, l, j) + jt(J, l, j, i, k) + jt(J, j, k, i, l)) / 4;
                    Q[(27 * l + 9 * k + 3 * i + j) * stride] += ci[m] * M2 + ci[m + 3] * H2;
                }
}

// doubleinner22(doubleinner42(Q, tau), v w) = sum Q_lkij tau_ji (v w)_kl
__device__ __forceinline__ double ortho_eps_vw(const double* Q, int stride, const double (&tau)[9], const double* v,
                                               const double* w) {
    double r = 0.0;
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double e = 0.0;
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) e += Q[(27 * l + 9 * k + 3 * i + j) * stride] * tau[3 * j + i];
            r += e * v[k] * w[l];
        }
    return r;
}

}  // namespace

// Eij of orthotropic grains (Sachs; src/enhancementfactors.f90:134-188): one
// lane per parcel, Q in the lane's LDS column
__global__ void __launch_bounds__(64) eij_ortho_kernel(OrthoArgs a) {
    __shared__ double lds[81 * 64];
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.P) return;
    double* Q = lds + threadIdx.x;
    OrthoDist q[3];
    const double* xs[3] = {a.x1, a.x2, a.x3};
    // v given (its n00 above 1e-8) or v = b x n (src/moments.f90:371-383)
    const bool derived = !a.x3 || !(a.x3[p] > 1e-8);
#pragma unroll
    for (int m = 0; m < 3; m++) {
        if (m == 2 && derived) break;
        double xr[15];
#pragma unroll
        for (int r = 0; r < 15; r++) xr[r] = r < a.nq ? xs[m][(int64_t)r * a.P + p] : 0.0;
        ortho_dist(a.A2, a.A4F, xr, q[m]);
    }
    if (derived) ortho_third(q[0], q[1], q[2]);
#pragma unroll
    for (int t = 0; t < 81; t++) Q[t * 64] = 0.0;
#pragma unroll
    for (int m = 0; m < 3; m++) {
        double J[6][6];
        ortho_joints(q, derived, m, J);
        ortho_Q_add(q, m, J, a.ci, Q, 64);
    }
    double e[9];
    if (a.frames) {
#pragma unroll
        for (int i = 0; i < 9; i++) e[i] = a.frames[p * 9 + i];
    } else {
#pragma unroll
        for (int i = 0; i < 9; i++) e[i] = (i % 4 == 0) ? 1.0 : 0.0;
    }
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
        const double num = ortho_eps_vw(Q, 64, tau, v, w);
        const double den = ortho_eps_vw(a.Qiso, 1, tau, v, w);   // the isotropic fabric (ortho_Q_kernel)
        a.Eij[p * 6 + s] = num / den;
    }
}

// Q of the isotropic fabric (q1 = q2 = q3 = the isotropic state; the
// reference's ai_orthotropic(qlm_iso, qlm_iso, qlm_iso)) into Qiso (81)
__global__ void ortho_Q_kernel(OrthoArgs a, double* Qiso) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    OrthoDist q[3];
    double xr[15];
#pragma unroll
    for (int r = 0; r < 15; r++) xr[r] = r == 0 ? 1.0 : 0.0;
#pragma unroll
    for (int m = 0; m < 3; m++) ortho_dist(a.A2, a.A4F, xr, q[m]);
    for (int t = 0; t < 81; t++) Qiso[t] = 0.0;
    for (int m = 0; m < 3; m++) {
        double J[6][6];
        ortho_joints(q, false, m, J);
        ortho_Q_add(q, m, J, a.ci, Qiso, 1);
    }
}

// elastic Vi of orthotropic grains (Voigt; src/wavepropagation.f90:41-68,
// 98-115): one lane per (parcel, direction); the distributions rotated as
// rotate_nlm4 rotates them, the acoustic tensor for k = z, Cardano
__global__ void __launch_bounds__(128) vi_ortho_kernel(OrthoArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.P * a.ndir) return;
    const int64_t p = t % a.P;
    const int d = (int)(t / a.P);
    const double* R = a.R + (int64_t)d * ROT4_STRIDE;
    OrthoDist q[3];
    const double* xs[3] = {a.x1, a.x2, a.x3};
    const bool derived = !a.x3 || !(a.x3[p] > 1e-8);   // n00 is invariant under rotation
#pragma unroll
    for (int m = 0; m < 3; m++) {
        if (m == 2 && derived) break;
        double xr[15];
        rotated_block(xs[m], a.P, p, a.nq, R, xr);
        ortho_dist(a.A2, a.A4F, xr, q[m]);
    }
    if (derived) ortho_third(q[0], q[1], q[2]);
    double Qn[9];
#pragma unroll
    for (int i = 0; i < 9; i++) Qn[i] = 0.0;
#pragma unroll
    for (int m = 0; m < 3; m++) {
        double J[6][6];
        ortho_joints(q, derived, m, J);
        const double* a2 = q[m].a2;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                // mu/2 (a2 + a2.kk + kk.a2 + (a2:kk) I) + lam a4:kk + lam' 2 sym2(J) ^{first-last} kk
                const double anti = (j == 2 ? a2[3 * i + 2] : 0.0) + (i == 2 ? a2[3 * 2 + j] : 0.0);
                double q4 = a4e(q[m], i, j, 2, 2);
                const double jf = jt(J, 2, i, j, 2) + jt(J, j, 2, 2, i);
                Qn[3 * i + j] += a.mu[m] / 2 * (a2[3 * i + j] + anti + a2[8] * (i == j ? 1.0 : 0.0)) + a.lam[m] * q4 +
                                 a.lam[m + 3] * jf;
            }
    }
    double vi[3];
    qnorm_to_vi(Qn, a.rho, vi);
    double* o = a.Vi + (p * a.ndir + d) * 3;
    o[0] = vi[0];
    o[1] = vi[1];
    o[2] = vi[2];
}

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

hipError_t launch_eij_ortho(const OrthoArgs& args, double* Qiso, hipStream_t s) {
    if (args.P == 0) return hipSuccess;
    hipLaunchKernelGGL(ortho_Q_kernel, dim3(1), dim3(64), 0, s, args, Qiso);
    OrthoArgs a = args;
    a.Qiso = Qiso;
    hipLaunchKernelGGL(eij_ortho_kernel, dim3(nblk(a.P, 64)), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_vi_ortho(const OrthoArgs& a, hipStream_t s) {
    if (a.P * a.ndir == 0) return hipSuccess;
    hipLaunchKernelGGL(vi_ortho_kernel, dim3(nblk(a.P * a.ndir, 128)), dim3(128), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_a4full(const A4FullArgs& a, hipStream_t s) {
    if (a.P == 0) return hipSuccess;
    hipLaunchKernelGGL(a4full_kernel, dim3(nblk(a.P, 128)), dim3(128), 0, s, a);
    return hipGetLastError();
}
#endif

}  // namespace sfg
