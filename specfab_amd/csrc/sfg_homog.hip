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
#include <cmath>

#include "generated/sfg_moment_maps.h"
#include "generated/sfg_moments68.h"
#include "sfg_a4slots.h"
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
    // descending order (the order of the swaps of a three-element sort);
    // selects instead of indexed reads keep everything in registers
    double d[3] = {a[0][0], a[1][1], a[2][2]};
    double w[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) w[i][k] = v[k][i];   // w[i] = eigenvector of d[i]
    auto swap_if = [&](int i, int j) {
        const bool sw = d[j] > d[i];
        const double di = d[i];
        d[i] = sw ? d[j] : di;
        d[j] = sw ? di : d[j];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double wi = w[i][k];
            w[i][k] = sw ? w[j][k] : wi;
            w[j][k] = sw ? wi : w[j][k];
        }
    };
    swap_if(0, 1);
    swap_if(0, 2);
    swap_if(1, 2);
#pragma unroll
    for (int i = 0; i < 3; i++) {
        lami[i] = d[i];
#pragma unroll
        for (int k = 0; k < 3; k++) ei[3 * i + k] = w[i][k];
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
__device__ inline void d_taylorP(const double* a2m, const double* a2v, const double (&a4v)[6][6], double tA,
                                 double tB, double tC, double Pm[6][6]) {
    const double s = sqrt(2.0);
    const double* a = a2m;
    const double Lm[6][6] = {
        {2 * a[0], 0, 0, 0, s * a[2], s * a[1]},
        {0, 2 * a[4], 0, s * a[5], 0, s * a[1]},
        {0, 0, 2 * a[8], s * a[5], s * a[2], 0},
        {0, s * a[5], s * a[5], a[4] + a[8], a[1], a[2]},
        {s * a[2], 0, s * a[2], a[1], a[0] + a[8], a[5]},
        {s * a[1], s * a[1], 0, a[2], a[5], a[0] + a[4]}};
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++)
            Pm[i][j] = (i == j ? 1.0 : 0.0) - tA * ((i < 3 ? 1.0 : 0.0) * a2v[j]) + tB * a4v[i][j] + tC * Lm[i][j];
}

// L L^T x = b with the packed factor of d_chol6 and the reciprocals of its
// diagonal (the forward and back substitutions multiply by them)
__device__ inline void d_chol6_solve_rcp(const double* Lf, const double* rd, double* b) {
#define LF(i, j) Lf[(i) * ((i) + 1) / 2 + (j)]
#pragma unroll
    for (int i = 0; i < 6; i++) {
        double s = b[i];
#pragma unroll
        for (int k = 0; k < i; k++) s -= LF(i, k) * b[k];
        b[i] = s * rd[i];
    }
#pragma unroll
    for (int i = 5; i >= 0; i--) {
        double s = b[i];
#pragma unroll
        for (int k = i + 1; k < 6; k++) s -= LF(k, i) * b[k];
        b[i] = s * rd[i];
    }
#undef LF
}

// a2 (Mandel, with the isotropic part) and a4 (Mandel 6x6) of parcel p from the
// l <= 4 rows of the state (src/moments.f90:164-178)
__device__ __forceinline__ void load_moments(const EijArgs& a, int64_t p, double (&a2v)[6], double (&a4v)[6][6]) {
    double xq[15];
#pragma unroll
    for (int q = 0; q < 15; q++) xq[q] = a.x[(int64_t)q * a.P + p];
    moment_maps(xq, a2v, a4v);
    const double ix0 = 1.0 / xq[0];
#pragma unroll
    for (int i = 0; i < 6; i++) a2v[i] = a2v[i] * ix0 + (i < 3 ? 1.0 / 3.0 : 0.0);
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j < 6; j++) a4v[i][j] *= ix0;
}

// Mandel vectors of one stress state q of the frame e (src/enhancementfactors.f90:
// 398-413): tv = mandel(tau) with tau = I/3 - v v (q < 3) or v w + w v, and
// uv = mandel(sym(v w)); returns v . w.  (v, w) = (e_q, e_q) for q < 3, else
// (e2, e3), (e1, e3), (e1, e2), picked by selects rather than indexed
// register reads.
__device__ __forceinline__ double stress_state(const double (&e)[9], int q, double (&tv)[6], double (&uv)[6]) {
    const double r2 = sqrt(2.0);
    const int iv = q < 3 ? q : (q == 3 ? 1 : 0), iw = q < 3 ? q : (q == 5 ? 1 : 2);
    double v[3], w[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        v[i] = iv == 0 ? e[i] : (iv == 1 ? e[3 + i] : e[6 + i]);
        w[i] = iw == 0 ? e[i] : (iw == 1 ? e[3 + i] : e[6 + i]);
    }
    const bool nrm = q < 3;
    tv[0] = nrm ? 1.0 / 3.0 - v[0] * v[0] : 2 * v[0] * w[0];
    tv[1] = nrm ? 1.0 / 3.0 - v[1] * v[1] : 2 * v[1] * w[1];
    tv[2] = nrm ? 1.0 / 3.0 - v[2] * v[2] : 2 * v[2] * w[2];
    tv[3] = r2 * (nrm ? -v[1] * v[2] : v[1] * w[2] + w[1] * v[2]);
    tv[4] = r2 * (nrm ? -v[0] * v[2] : v[0] * w[2] + w[0] * v[2]);
    tv[5] = r2 * (nrm ? -v[0] * v[1] : v[0] * w[1] + w[0] * v[1]);
    uv[0] = v[0] * w[0];
    uv[1] = v[1] * w[1];
    uv[2] = v[2] * w[2];
    uv[3] = r2 * (0.5 * (v[1] * w[2] + w[1] * v[2]));
    uv[4] = r2 * (0.5 * (v[0] * w[2] + w[0] * v[2]));
    uv[5] = r2 * (0.5 * (v[0] * w[1] + w[0] * v[1]));
    return v[0] * w[0] + v[1] * w[1] + v[2] * w[2];
}

// lower triangles (packed, (i, j) at i (i + 1) / 2 + j, j <= i) of the Taylor
// matrix P = I - tA 1 (x) a2v + tB a4v + tC L(a2) (src/homogenizations.f90:
// 164-172; dposv('L') reads only this triangle) and of the symmetric Sachs
// matrix Q = sB a4v + sC L(a2), with L the Mandel anticommutator
// (src/homogenizations.f90:238-258)
__device__ __forceinline__ void taylor_sachs_lower(const EijArgs& c, const double (&a2v)[6],
                                                   const double (&a4v)[6][6], double (&Pl)[21], double (&Ql)[21]) {
    const double s = sqrt(2.0), r2i = 0.70710678118654752440;
    const double m[9] = {a2v[0], a2v[5] * r2i, a2v[4] * r2i, a2v[5] * r2i, a2v[1], a2v[3] * r2i,
                         a2v[4] * r2i, a2v[3] * r2i, a2v[2]};
    const double Lm[6][6] = {
        {2 * m[0], 0, 0, 0, s * m[2], s * m[1]},
        {0, 2 * m[4], 0, s * m[5], 0, s * m[1]},
        {0, 0, 2 * m[8], s * m[5], s * m[2], 0},
        {0, s * m[5], s * m[5], m[4] + m[8], m[1], m[2]},
        {s * m[2], 0, s * m[2], m[1], m[0] + m[8], m[5]},
        {s * m[1], s * m[1], 0, m[2], m[5], m[0] + m[4]}};
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = 0; j <= i; j++) {
            Pl[i * (i + 1) / 2 + j] =
                (i == j ? 1.0 : 0.0) - c.tA * ((i < 3 ? 1.0 : 0.0) * a2v[j]) + c.tB * a4v[i][j] + c.tC * Lm[i][j];
            Ql[i * (i + 1) / 2 + j] = c.sB * a4v[i][j] + c.sC * Lm[i][j];
        }
}

// in-place Cholesky of a packed lower triangle, the arithmetic of d_chol6
__device__ inline bool d_chol6_packed(double (&F)[21]) {
#define PF(i, j) F[(i) * ((i) + 1) / 2 + (j)]
#pragma unroll
    for (int j = 0; j < 6; j++) {
        double d = PF(j, j);
#pragma unroll
        for (int k = 0; k < j; k++) d -= PF(j, k) * PF(j, k);
        if (!(d > 0)) return false;
        d = sqrt(d);
        PF(j, j) = d;
        const double r = 1.0 / d;
#pragma unroll
        for (int i = j + 1; i < 6; i++) {
            double t = PF(i, j);
#pragma unroll
            for (int k = 0; k < j; k++) t -= PF(i, k) * PF(j, k);
            PF(i, j) = t * r;
        }
    }
    return true;
#undef PF
}

// the Sachs ratio es : vw / (siso tau : vw) of one stress state
__device__ __forceinline__ double sachs_ratio(const EijArgs& c, const double (&a2v)[6], const double (&Ql)[21],
                                              const double (&tv)[6], const double (&uv)[6], double vdw,
                                              double* tvw_out) {
    double tvw = 0, I2 = 0, a2t = 0, qt = 0;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        tvw += tv[i] * uv[i];
        I2 += tv[i] * tv[i];
        a2t += a2v[i] * tv[i];
        double r = 0;
#pragma unroll
        for (int j = 0; j < 6; j++) {
            const int hi = i > j ? i : j, lo = i > j ? j : i;
            r += Ql[hi * (hi + 1) / 2 + lo] * tv[j];
        }
        qt += uv[i] * r;
    }
    const bool lin = c.n_grain == 1;
    const double esvw = (lin ? 1.0 : I2) * (tvw - c.sA * a2t * vdw + qt);
    *tvw_out = tvw;
    return esvw / ((lin ? c.sachs_iso : I2) * tvw);
}

// a stress state given directly (st = v(3), w(3), tau(9) row-major, tau
// symmetric): tv = mandel(tau), uv = mandel(sym(v w)); returns v . w
__device__ __forceinline__ double given_state(const double* st, double (&tv)[6], double (&uv)[6]) {
    const double r2 = sqrt(2.0);
    const double *v = st, *w = st + 3, *t = st + 6;
    tv[0] = t[0];
    tv[1] = t[4];
    tv[2] = t[8];
    tv[3] = r2 * t[5];
    tv[4] = r2 * t[2];
    tv[5] = r2 * t[1];
    uv[0] = v[0] * w[0];
    uv[1] = v[1] * w[1];
    uv[2] = v[2] * w[2];
    uv[3] = r2 * (0.5 * (v[1] * w[2] + w[1] * v[2]));
    uv[4] = r2 * (0.5 * (v[0] * w[2] + w[0] * v[2]));
    uv[5] = r2 * (0.5 * (v[0] * w[1] + w[0] * v[1]));
    return v[0] * w[0] + v[1] * w[1] + v[2] * w[2];
}

// Evw of a parcel whose Taylor matrix P is not positive definite: the
// reference's regularized branch (src/homogenizations.f90:174-185), solving
// (F^T F + 1e-6 I) x = F^T tau with F the matrix its failed dposv('L') left
// (the REAL(4) 1e-6), for the nq stress states state(q, tv, uv) (returning
// v . w).  Rare: the Eij path runs it in a second pass
// (eij_regularized_kernel) over the rows the first one marked, so that its
// registers do not weigh on the common path.
template <class SF>
__device__ __noinline__ int d_states_regularized(const EijArgs& c, int64_t p, int nq, SF state, double* Eij) {
    int status = 2;
    double a2v[6], a4v[6][6];
    load_moments(c, p, a2v, a4v);
    const double r2i = 0.70710678118654752440;
    const double a2m[9] = {a2v[0], a2v[5] * r2i, a2v[4] * r2i, a2v[5] * r2i, a2v[1], a2v[3] * r2i,
                           a2v[4] * r2i, a2v[3] * r2i, a2v[2]};
    double Pl[21], Ql[21], F[6][6], R[6][6], Lf[21];
    taylor_sachs_lower(c, a2v, a4v, Pl, Ql);
    d_taylorP(a2m, a2v, a4v, c.tA, c.tB, c.tC, F);
    d_failed_factor6(F);
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            double t = 0;
            for (int k = 0; k < 6; k++) t += F[k][i] * F[k][j];
            R[i][j] = t + (i == j ? (double)1e-6f : 0.0);
        }
    const bool ok_reg = d_chol6(R, Lf);
    if (!ok_reg) status |= 4;
    for (int q = 0; q < nq; q++) {
        double tv[6], uv[6], tvw;
        const double vdw = state(q, tv, uv);
        const double ts = sachs_ratio(c, a2v, Ql, tv, uv, vdw, &tvw);
        double x[6];
        for (int i = 0; i < 6; i++) {
            x[i] = 0;
            for (int k = 0; k < 6; k++) x[i] += F[k][i] * tv[k];
        }
        if (ok_reg) d_chol6_solve(Lf, x);
        double etvw = 0;
        for (int i = 0; i < 6; i++) etvw += x[i] * uv[i];
        const double tt = etvw / (c.taylor_iso * tvw);
        const double E = (1 - c.alpha) * ts + c.alpha * tt;
        Eij[q] = E;
        if (!isfinite(E)) status |= 1;
    }
    return status;
}

__device__ __forceinline__ int d_Eij_regularized(const EijArgs& c, int64_t p, const double (&e)[9], double* Eij) {
    return d_states_regularized(
        c, p, 6, [&](int q, double (&tv)[6], double (&uv)[6]) { return stress_state(e, q, tv, uv); }, Eij);
}

// Eij of one parcel in frame (e1, e2, e3): src/enhancementfactors.f90:23-69,
// with every contraction written in Mandel 6-vectors.  For a stress state tau
// (tv = mandel(tau)) and the projector vw = v (x) w (uv = mandel(sym(vw))):
//   tau : vw = tv . uv,   I : vw = v . w,   (a4 : tau) : vw = uv . a4v tv,
//   (tau a2 + a2 tau) : vw = uv . L(a2) tv,
// so the Sachs term es : vw = f [tv.uv - sA (a2v.tv)(v.w) + uv.Q tv] with
// Q = sB a4v + sC L(a2) formed once per parcel and f = 1 (n' = 1) or I2 =
// tau : tau (n' = -3) (src/homogenizations.f90:69-116), and the Taylor term
// et : vw = (P^-1 tv) . uv (:145-189).  Pl: the packed lower triangle of P,
// factored here in place.
// marks a row whose Taylor matrix is not positive definite for the second
// pass (a NaN payload that arithmetic does not produce)
constexpr unsigned long long kRegularizeMark = 0x7ff8f00dfeedc0deull;

__device__ int d_Eij(const EijArgs& c, int64_t p, const double (&a2v)[6], double (&Pl)[21], const double (&Ql)[21],
                     const double (&e)[9], double* Eij) {
    if (!d_chol6_packed(Pl)) {
        Eij[0] = __longlong_as_double((long long)kRegularizeMark);
        return 2;
    }
    double rd[6];
#pragma unroll
    for (int i = 0; i < 6; i++) rd[i] = 1.0 / Pl[i * (i + 1) / 2 + i];
    int status = 0;
#pragma unroll 1
    for (int q = 0; q < 6; q++) {
        double tv[6], uv[6], tvw;
        const double vdw = stress_state(e, q, tv, uv);
        const double ts = sachs_ratio(c, a2v, Ql, tv, uv, vdw, &tvw);
        double xs[6];
#pragma unroll
        for (int i = 0; i < 6; i++) xs[i] = tv[i];
        d_chol6_solve_rcp(Pl, rd, xs);
        double etvw = 0;
#pragma unroll
        for (int i = 0; i < 6; i++) etvw += xs[i] * uv[i];
        const double tt = etvw / (c.taylor_iso * tvw);
        const double E = (1 - c.alpha) * ts + c.alpha * tt;
        Eij[q] = E;
        if (!isfinite(E)) status |= 1;
    }
    return status;
}

// the frame (e1, e2, e3) of parcel p: its a2 eigenframe (eigenvalues to lam),
// the caller's frame or the Cartesian axes
__device__ __forceinline__ void eij_frame(const EijArgs& a, int64_t p, const double (&a2v)[6], double (&e)[9],
                                          double (&lam)[3]) {
    lam[0] = lam[1] = lam[2] = 0.0;
    if (a.frame == 0) {
        const double r2i = 0.70710678118654752440;
        const double a2m[9] = {a2v[0], a2v[5] * r2i, a2v[4] * r2i, a2v[5] * r2i, a2v[1], a2v[3] * r2i,
                               a2v[4] * r2i, a2v[3] * r2i, a2v[2]};
        d_eig3(a2m, e, lam);
    } else if (a.frame == 1) {
#pragma unroll
        for (int i = 0; i < 9; i++) e[i] = a.frames[p * 9 + i];
    } else {
#pragma unroll
        for (int i = 0; i < 9; i++) e[i] = (i % 4 == 0) ? 1.0 : 0.0;
    }
}

// returns the row's status bits (2: marked for the regularized second pass)
__device__ __forceinline__ int eij_item(const EijArgs& a, int64_t t) {
    const int64_t p = a.idx ? (int64_t)a.idx[t] : t;   // state column
    const int64_t o = t;                               // output row
    double a2v[6], Pl[21], Ql[21];
    {
        double a4v[6][6];
        load_moments(a, p, a2v, a4v);
        taylor_sachs_lower(a, a2v, a4v, Pl, Ql);   // a4 is not needed after this
    }
    double e[9], lam[3];
    eij_frame(a, p, a2v, e, lam);
    const int st = d_Eij(a, p, a2v, Pl, Ql, e, a.Eij + o * 6);
    if (a.lami)
#pragma unroll
        for (int i = 0; i < 3; i++) a.lami[o * 3 + i] = lam[i];
    if (a.ei)
#pragma unroll
        for (int i = 0; i < 9; i++) a.ei[o * 9 + i] = e[i];
    if (a.status && st) atomicOr(a.status + p, st);
    return st;
}

// two waves per SIMD (at most 256 registers per lane); with wave flags every
// wavefront writes its own flag, so they need no clearing
// eigenpairs of P symmetric 3x3 matrices M (P, 9): ei (P, 9) rows e1, e2, e3,
// lami (P, 3), largest first (eig3, src/frames.f90:62-80)
__global__ void __launch_bounds__(256) eig3_kernel(int64_t P, const double* M, double* ei, double* lami) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    double m[9], e[9], l[3];
#pragma unroll
    for (int i = 0; i < 9; i++) m[i] = M[p * 9 + i];
    d_eig3(m, e, l);
#pragma unroll
    for (int i = 0; i < 9; i++) ei[p * 9 + i] = e[i];
#pragma unroll
    for (int i = 0; i < 3; i++) lami[p * 3 + i] = l[i];
}

__global__ void __launch_bounds__(128, 2) eij_kernel(EijArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = a.idx ? a.nidx : a.P;
    int st = 0;
    if (t < n) st = eij_item(a, t);
    if (a.wave_flag) {
        const unsigned long long m = __ballot((st & 2) != 0);
        if ((threadIdx.x & 63) == 0 && t < n) a.wave_flag[t >> 6] = m != 0;
    }
}

// work list whose length is on the device: grid-stride loop (a separate
// kernel, so that the loop does not cost the one-item kernel registers)
__global__ void __launch_bounds__(128, 2) eij_list_kernel(EijArgs a) {
    const int64_t n = *a.nidx_dev < a.nidx ? *a.nidx_dev : a.nidx;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) eij_item(a, t);
}

// second pass: the rows the first pass marked (Taylor matrix not positive
// definite), with the reference's regularized branch
__device__ __forceinline__ void eij_regularized_item(const EijArgs& a, int64_t t) {
    double* out = a.Eij + t * 6;
    if ((unsigned long long)__double_as_longlong(out[0]) != kRegularizeMark) return;
    const int64_t p = a.idx ? (int64_t)a.idx[t] : t;
    double a2v[6], a4v[6][6], e[9], lam[3];
    load_moments(a, p, a2v, a4v);
    eij_frame(a, p, a2v, e, lam);
    const int st = d_Eij_regularized(a, p, e, out);
    if (a.status) atomicOr(a.status + p, st);
}

// the 64 rows of each flagged wavefront (batch form with wave flags)
__global__ void __launch_bounds__(256) eij_regularized_waves_kernel(EijArgs a) {
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= (a.P + 63) / 64 || !a.wave_flag[w]) return;
    const int64_t end = (w + 1) * 64 < a.P ? (w + 1) * 64 : a.P;
    for (int64_t t = w * 64; t < end; t++) eij_regularized_item(a, t);
}

// every row (lists, or no wave flags)
__global__ void __launch_bounds__(256) eij_regularized_kernel(EijArgs a) {
    const int64_t n = (a.idx && a.nidx_dev) ? (*a.nidx_dev < a.nidx ? *a.nidx_dev : a.nidx) : (a.idx ? a.nidx : a.P);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) eij_regularized_item(a, t);
}

// Evw of the given stress states per parcel (src/enhancementfactors.f90:47-69):
// one lane per parcel, the Taylor factor formed once for all states; a
// parcel whose Taylor matrix is not positive definite takes the regularized
// branch in place (rare)
__global__ void __launch_bounds__(128, 2) evw_kernel(EijArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.P) return;
    double a2v[6], Pl[21], Ql[21];
    {
        double a4v[6][6];
        load_moments(a, p, a2v, a4v);
        taylor_sachs_lower(a, a2v, a4v, Pl, Ql);
    }
    double* out = a.Eij + p * a.nstates;
    const double* S = a.states;
    int st = 0;
    if (!d_chol6_packed(Pl)) {
        st = d_states_regularized(
            a, p, a.nstates, [&](int q, double (&tv)[6], double (&uv)[6]) { return given_state(S + 15 * q, tv, uv); },
            out);
    } else {
        double rd[6];
#pragma unroll
        for (int i = 0; i < 6; i++) rd[i] = 1.0 / Pl[i * (i + 1) / 2 + i];
#pragma unroll 1
        for (int q = 0; q < a.nstates; q++) {
            double tv[6], uv[6], tvw;
            const double vdw = given_state(S + 15 * q, tv, uv);
            const double ts = sachs_ratio(a, a2v, Ql, tv, uv, vdw, &tvw);
            double xs[6];
#pragma unroll
            for (int i = 0; i < 6; i++) xs[i] = tv[i];
            d_chol6_solve_rcp(Pl, rd, xs);
            double etvw = 0;
#pragma unroll
            for (int i = 0; i < 6; i++) etvw += xs[i] * uv[i];
            const double tt = etvw / (a.taylor_iso * tvw);
            const double E = (1 - a.alpha) * ts + a.alpha * tt;
            out[q] = E;
            if (!isfinite(E)) st |= 1;
        }
    }
    if (a.status && st) atomicOr(a.status + p, st);
}


// ---------------------------------------------------------------------------
// n' = 3 Sachs homogenization (src/homogenizations.f90:93-103) and the
// basal-RSS deformabilities <D> (n_RSS = 2, 4; src/dynamics.f90:385-449).
//
// The fluidity of an n' = 3 grain makes the grain average an expectation
// of polynomials in the c-axis of degree <= 8.  a2 and a4 enter as the
// reference's structure-tensor bodies (the full-a4 body with its REAL(4)
// literals and its a4(3,2,1,2) entry); a6 and a8 only through products with
// tau, evaluated as <c_i c_j P(c)> with P a polynomial in c built from the
// quadratic forms c.tau.c and c.tau^2.c, against the exact sixth and eighth
// moments (tables/make_moment_maps.py; the reference's a6 / a8 bodies carry
// REAL(4) literal coefficients, so those terms agree with it to ~1e-7).

__constant__ double kMOM8[45][45] = SFG_MOM8_INIT;
__constant__ double kMOM6_ISO[28] = SFG_MOM6_ISO_INIT;
__constant__ double kMOM8_ISO[45] = SFG_MOM8_ISO_INIT;
__constant__ double kMOM6[28][28] = SFG_MOM6_INIT;

namespace {

// element k of a lane's column in LDS (stride 64: one column per lane of the
// wavefront) or of a plain array (stride 1)
template <int S>
struct Col {
    const double* p;
    __device__ __forceinline__ double operator[](int k) const { return p[k * S]; }
};

// index of x^a y^b z^(d-a-b) among the degree-d monomials (a = d..0, b = d-a..0)
__host__ __device__ constexpr int midx(int d, int a, int b) { return (d - a) * (d - a + 1) / 2 + (d - a - b); }

// c.S.c for symmetric S (row-major) as a degree-2 polynomial
__device__ __forceinline__ void quad_form(const double (&S)[9], double (&P)[6]) {
    P[0] = S[0], P[1] = 2 * S[1], P[2] = 2 * S[2], P[3] = S[4], P[4] = 2 * S[5], P[5] = S[8];
}

// C = A B for homogeneous polynomials of degrees DA and DB
template <int DA, int DB>
__device__ __forceinline__ void poly_mul(const double* A, const double* B, double* C) {
    constexpr int D = DA + DB, NC = (D + 1) * (D + 2) / 2;
#pragma unroll
    for (int k = 0; k < NC; k++) C[k] = 0.0;
#pragma unroll
    for (int a1 = DA; a1 >= 0; a1--)
#pragma unroll
        for (int b1 = DA - a1; b1 >= 0; b1--)
#pragma unroll
            for (int a2 = DB; a2 >= 0; a2--)
#pragma unroll
                for (int b2 = DB - a2; b2 >= 0; b2--)
                    C[midx(D, a1 + a2, b1 + b2)] += A[midx(DA, a1, b1)] * B[midx(DB, a2, b2)];
}

// sum_kl a4(i,j,k,l) S_kl
template <class U>
__device__ __forceinline__ double a4_dot(const U& u, int i, int j, const double (&S)[9]) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int l = 0; l < 3; l++) s += u[kA4Slots.s[27 * i + 9 * j + 3 * k + l]] * S[3 * k + l];
    return s;
}

// the reference's full a2 (f_ev_c2 body) of the l <= 2 rows
__device__ __forceinline__ void a2_full(const double* xq, double (&a2)[9]) {
    const double n00 = xq[0];
    const double h0 = xq[1] / n00, h1r = xq[2] / n00, h1i = xq[3] / n00, h2r = xq[4] / n00, h2i = xq[5] / n00;
    const double c = sqrt(2.0 / 3), s = sqrt(2 / 15.0);
    double e[9];
    e[0] = -0.5 * c * h0 + h2r;
    e[4] = -0.5 * c * h0 - h2r;
    e[8] = c * h0;
    e[1] = e[3] = -h2i;
    e[2] = e[6] = -h1r;
    e[5] = e[7] = h1i;
#pragma unroll
    for (int i = 0; i < 9; i++) a2[i] = s * e[i] + ((i % 4 == 0) ? 1.0 / 3.0 : 0.0);
}

// <A B> for homogeneous polynomials A (degree DA) and B (degree DB) against
// the moments m of degree DA + DB
template <int DA, int DB, class M>
__device__ __forceinline__ double pdot(const double* A, const double* B, const M& m) {
    double s = 0.0;
#pragma unroll
    for (int a1 = DA; a1 >= 0; a1--)
#pragma unroll
        for (int b1 = DA - a1; b1 >= 0; b1--) {
            double t = 0.0;
#pragma unroll
            for (int a2 = DB; a2 >= 0; a2--)
#pragma unroll
                for (int b2 = DB - a2; b2 >= 0; b2--)
                    t += B[midx(DB, a2, b2)] * m[midx(DA + DB, a1 + a2, b1 + b2)];
            s += A[midx(DA, a1, b1)] * t;
        }
    return s;
}

// c.X.c for a general X (its symmetric part) as a degree-2 polynomial
__device__ __forceinline__ void quad_form_gen(const double (&X)[9], double (&P)[6]) {
    P[0] = X[0], P[1] = X[1] + X[3], P[2] = X[2] + X[6], P[3] = X[4], P[4] = X[5] + X[7], P[5] = X[8];
}

// eps(tau) : vw of the n' = 3 Sachs average (src/homogenizations.f90:93-103,
// 115), as the reference contracts it: doubleinner22(eps, v w) = sum_ij
// eps_ij w_i v_j (eps is not symmetric: a4's (3,2,1,2) entry).  With
// Y = w v^T, Z = tau Y + Y tau and E2 = <eta c c>:
//   eps : Y = c0 (tau : Y) - A (E2 : tau) tr Y + B (A4t : Y) + C (E2 : Z)
// where <eta> = c0, E2 : X = I2 a2:X + B <(cXc) (c tau c)^2> + 2C a4[X]:tau^2,
// A4t : Y = I2 a4[Y]:tau + B <(cYc) (c tau c)^3> + 2C <(cYc) (c tau c)(c tau^2 c)>
template <class U, class M6, class M8>
__device__ __forceinline__ double sachs3_vw(const double (&tau)[9], const double (&v)[3], const double (&w)[3],
                                            const double (&a2)[9], const U& u, const M6& m6, const M8& m8,
                                            double sA, double sB, double sC) {
    double t2[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            t2[3 * i + j] = tau[3 * i] * tau[j] + tau[3 * i + 1] * tau[3 + j] + tau[3 * i + 2] * tau[6 + j];
    double I2 = 0.0;
#pragma unroll
    for (int i = 0; i < 9; i++) I2 += tau[i] * tau[i];
    double P1[6], P3[6], P11[15];
    quad_form(tau, P1);
    quad_form(t2, P3);
    poly_mul<2, 2>(P1, P1, P11);
    double Y[9], Z[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) Y[3 * i + j] = w[i] * v[j];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            double z = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) z += tau[3 * i + k] * Y[3 * k + j] + Y[3 * i + k] * tau[3 * k + j];
            Z[3 * i + j] = z;
        }
    const double trY = v[0] * w[0] + v[1] * w[1] + v[2] * w[2];
    // <eta>, tau : Y, a4[Y] : tau and the a2 / a4 parts of E2 : tau and E2 : Z
    double c0 = 0.0, a2t2 = 0.0, tY = 0.0, a4Yt = 0.0, E2t = 0.0, E2Z = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int ij = 3 * i + j;
            const double a4t = a4_dot(u, i, j, tau), a4t2 = a4_dot(u, i, j, t2);
            c0 += a4t * tau[ij];
            a2t2 += a2[ij] * t2[ij];
            tY += tau[ij] * Y[ij];
            a4Yt += Y[ij] * a4t;
            E2t += tau[ij] * (I2 * a2[ij] + 2 * sC * a4t2);
            E2Z += Z[ij] * (I2 * a2[ij] + 2 * sC * a4t2);
        }
    c0 = I2 + sB * c0 + 2 * sC * a2t2;
    double PX[6];
    quad_form(tau, PX);
    E2t += sB * pdot<2, 4>(PX, P11, m6);
    quad_form_gen(Z, PX);
    E2Z += sB * pdot<2, 4>(PX, P11, m6);
    double R[15];
    quad_form_gen(Y, PX);
    poly_mul<2, 2>(PX, P1, R);
    const double A4tY = I2 * a4Yt + sB * pdot<4, 4>(R, P11, m8) + 2 * sC * pdot<4, 2>(R, P3, m6);
    return c0 * tY - sA * E2t * trY + sB * A4tY + sC * E2Z;
}

// ev_D4 (src/dynamics.f90:424-449): 35/2 <RSS^4> / (tau:tau)^2 with
// RSS^2 = c.tau^2.c - (c.tau.c)^2
template <class U, class M6, class M8>
__device__ __forceinline__ double ev_D4_of(const double (&tau)[9], const U& u, const M6& m6, const M8& m8) {
    double t2[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            t2[3 * i + j] = tau[3 * i] * tau[j] + tau[3 * i + 1] * tau[3 + j] + tau[3 * i + 2] * tau[6 + j];
    const double norm = t2[0] + t2[4] + t2[8];
    double d = 0.0;                                          // tau^2 : a4 : tau^2
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) d += t2[3 * i + j] * a4_dot(u, i, j, t2);
    double P1[6], P3[6], P11[15];
    quad_form(tau, P1);
    quad_form(t2, P3);
    poly_mul<2, 2>(P1, P1, P11);
    const double q4 = pdot<4, 4>(P11, P11, m8);   // <(c tau c)^4>
    const double q6 = pdot<4, 2>(P11, P3, m6);    // <(c tau c)^2 (c tau^2 c)>
    d = d + q4 - 2 * q6;
    return 35 / 2.0 * d / (norm * norm);
}

// degree-6 moment k as the sum of three eighth moments (x^2 + y^2 + z^2 = 1)
struct M6Idx {
    int i[28][3];
};
__host__ __device__ constexpr M6Idx make_m6idx() {
    M6Idx t{};
    int k = 0;
    for (int a = 6; a >= 0; a--)
        for (int b = 6 - a; b >= 0; b--, k++) {
            t.i[k][0] = midx(8, a + 2, b);
            t.i[k][1] = midx(8, a, b + 2);
            t.i[k][2] = midx(8, a, b);
        }
    return t;
}
constexpr M6Idx kM6Idx = make_m6idx();

// the sixth moments of a lane's eighth-moment column in LDS
struct M6From8 {
    const double* p;
    __device__ __forceinline__ double operator[](int k) const {
        return p[kM6Idx.i[k][0] * 64] + p[kM6Idx.i[k][1] * 64] + p[kM6Idx.i[k][2] * 64];
    }
};

// a parcel's l <= 8 rows (missing rows zero) into xq, and its sorted full-a4
// entries and eighth moments into the lane's LDS columns (lds[k * 64]: 15 a4
// entries, then 45 eighth moments; 60 doubles per lane keep four 64-lane
// workgroups per CU; the sixth moments follow from the eighth)
constexpr int kColU = 0, kCol8 = 15, kCols = 60;
__device__ __forceinline__ void nonlin_moments(const double* x, int64_t P, int64_t p, int nrows, const double* A4F,
                                               double (&xq)[45], double* lds) {
#pragma unroll
    for (int q = 0; q < 45; q++) xq[q] = q < nrows ? x[(int64_t)q * P + p] : 0.0;
    const double ix0 = 1.0 / xq[0];
#pragma unroll 1
    for (int e = 0; e < 15; e++) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 15; q++) s += A4F[e * 15 + q] * xq[q];
        lds[(kColU + e) * 64] = s * ix0;
    }
#pragma unroll 1
    for (int k = 0; k < 45; k++) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 45; q++) s += kMOM8[k][q] * xq[q];
        lds[(kCol8 + k) * 64] = s * ix0;
    }
}

}  // namespace

// Eij with n' = 3 Sachs: (1 - alpha) Sachs + alpha Taylor, the Taylor values
// (alpha > 0) already in Eij from eij_kernel; writes the frame as eij_kernel.
// STATES: Evw of the given stress states (a.states) instead of the frame's six
template <bool STATES>
__global__ void __launch_bounds__(64) eij_sachs3_kernel(EijArgs a, Sachs3Args b) {
    __shared__ double lds[kCols * 64];
    __shared__ double iso[88];   // the isotropic state's a4 entries, sixth and eighth moments
    for (int k = threadIdx.x; k < 88; k += blockDim.x)
        iso[k] = k < 15 ? b.u_iso[k] : (k < 43 ? kMOM6_ISO[k - 15] : kMOM8_ISO[k - 43]);
    __syncthreads();
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.P) return;
    double* col = lds + threadIdx.x;
    double xq[45];
    nonlin_moments(a.x, a.P, p, b.nrows, b.A4F, xq, col);
    const Col<64> u{col + kColU * 64}, m8{col + kCol8 * 64};
    const M6From8 m6{col + kCol8 * 64};
    const Col<1> ui{iso}, m6i{iso + 15}, m8i{iso + 43};
    double a2v[6];
    {
        double xl[15], a4s[6][6];
#pragma unroll
        for (int q = 0; q < 15; q++) xl[q] = xq[q];
        moment_maps(xl, a2v, a4s);
        const double ix0 = 1.0 / xq[0];
#pragma unroll
        for (int i = 0; i < 6; i++) a2v[i] = a2v[i] * ix0 + (i < 3 ? 1.0 / 3.0 : 0.0);
    }
    double e[9], lam[3];
    if (!STATES) eij_frame(a, p, a2v, e, lam);
    double a2[9], a2i[9];
    a2_full(xq, a2);
#pragma unroll
    for (int i = 0; i < 9; i++) a2i[i] = (i % 4 == 0) ? 1.0 / 3.0 : 0.0;
    // the stress state q of the frame: tau_vv = I/3 - v v (q < 3), tau_vw = v w + w v
    // (src/enhancementfactors.f90:398-413)
    auto state = [&](int q, double (&v)[3], double (&w)[3], double (&tau)[9]) {
        if (STATES) {
            const double* st = a.states + 15 * q;
#pragma unroll
            for (int i = 0; i < 3; i++) v[i] = st[i], w[i] = st[3 + i];
#pragma unroll
            for (int i = 0; i < 9; i++) tau[i] = st[6 + i];
            return;
        }
        const int iv = q < 3 ? q : (q == 3 ? 1 : 0), iw = q < 3 ? q : (q == 5 ? 1 : 2);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            v[i] = iv == 0 ? e[i] : (iv == 1 ? e[3 + i] : e[6 + i]);
            w[i] = iw == 0 ? e[i] : (iw == 1 ? e[3 + i] : e[6 + i]);
        }
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                tau[3 * i + j] = q < 3 ? (i == j ? 1.0 / 3.0 : 0.0) - v[i] * v[j] : v[i] * w[j] + w[i] * v[j];
    };
    int st = 0;
    const int nq = STATES ? a.nstates : 6;
#pragma unroll 1
    for (int q = 0; q < nq; q++) {
        __asm__ volatile("" ::: "memory");   // reread the shared isotropic values in each pass
        double v[3], w[3], tau[9];
        state(q, v, w, tau);
        const double den = sachs3_vw(tau, v, w, a2i, ui, m6i, m8i, a.sA, a.sB, a.sC);
        const double num = sachs3_vw(tau, v, w, a2, u, m6, m8, a.sA, a.sB, a.sC);
        const double es = num / den;
        double* out = a.Eij + p * nq + q;
        const double E = a.alpha > 0 ? (1 - a.alpha) * es + a.alpha * *out : es;
        *out = E;
        if (!isfinite(E)) st |= 1;
    }
    if (!STATES && a.lami)
#pragma unroll
        for (int i = 0; i < 3; i++) a.lami[p * 3 + i] = lam[i];
    if (!STATES && a.ei)
#pragma unroll
        for (int i = 0; i < 9; i++) a.ei[p * 9 + i] = e[i];
    if (a.status && st) atomicOr(a.status + p, st);
}

// the exact sixth and eighth moments <x^a y^b z^c> (monomials in the order
// a = d..0, b = d - a..0) of P states: m6 (P, 28), m8 (P, 45)
__global__ void __launch_bounds__(128) moments68_kernel(int64_t P, int nrows, const double* x, double* m6,
                                                        double* m8) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    double xq[45];
#pragma unroll
    for (int q = 0; q < 45; q++) xq[q] = q < nrows ? x[(int64_t)q * P + p] : 0.0;
    const double ix0 = 1.0 / xq[0];
#pragma unroll 1
    for (int k = 0; k < 28; k++) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 28; q++) s += kMOM6[k][q] * xq[q];
        m6[p * 28 + k] = s * ix0;
    }
#pragma unroll 1
    for (int k = 0; k < 45; k++) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 45; q++) s += kMOM8[k][q] * xq[q];
        m8[p * 45 + k] = s * ix0;
    }
}

// <D> with n_RSS = 4 per parcel (tau (P, 3, 3)), optionally mapped to the
// CAFFE enhancement factor (caffe_kernel)
__global__ void __launch_bounds__(64) ev_D4_kernel(int64_t P, int nrows, const double* A4F, const double* x,
                                                   const double* tau, double* out) {
    __shared__ double lds[kCols * 64];
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    double* col = lds + threadIdx.x;
    double xq[45];
    nonlin_moments(x, P, p, nrows, A4F, xq, col);
    const Col<64> u{col + kColU * 64}, m8{col + kCol8 * 64};
    const M6From8 m6{col + kCol8 * 64};
    double t[9];
#pragma unroll
    for (int i = 0; i < 9; i++) t[i] = tau[p * 9 + i];
    out[p] = ev_D4_of(t, u, m6, m8);
}

// E_CAFFE (src/enhancementfactors.f90:302-332) from <D> in place
__global__ void __launch_bounds__(256) caffe_kernel(int64_t P, int n_rss, double Emin, double Emax, double* DE) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const double D = DE[p];
    const double Dmax = n_rss == 2 ? 5 / 2.0 : 35 / 8.0;
    const double Dmaxpow = pow(Dmax, 4.0 / n_rss);
    double E;
    if (D < 1) {
        const double gam = (4.0 / n_rss) / Dmaxpow * (Emax - 1) / (1 - Emin);
        E = Emin + (1 - Emin) * pow(D, gam);
    } else {
        E = ((Emax - 1) * pow(D, 4.0 / n_rss) + Dmaxpow - Emax) / (Dmaxpow - 1);
    }
    DE[p] = E;
}

// rheo_params_tranisotropic (src/rheologies.f90:123-132), d = 3, of the Sachs
// (ef = +1) and Taylor (ef = -1) branches and the isotropic normalisations
// (src/homogenizations.f90:207, 235): the same for every parcel, so on the host
static void eij_params(EijArgs& a) {
    auto rheo = [&](int ef, double* cI, double* cM, double* cL) {
        const double ne = (double)(ef * 2) / ((double)a.n_grain + 1);
        const double pc = std::pow(a.Ecc, ne), pa = std::pow(a.Eca, ne);
        *cI = (pc - 1) / 2;
        *cM = (3 * (pc + 1) - 2) / 2 - 2 * pa;
        *cL = pa - 1;
    };
    rheo(1, &a.sA, &a.sB, &a.sC);
    rheo(-1, &a.tA, &a.tB, &a.tC);
    a.sachs_iso = 1 + 2.0 / 15 * a.sB + 2.0 / 3 * a.sC;            // src/homogenizations.f90:207
    a.taylor_iso = 1.0 / (1 + 2.0 / 15 * a.tB + 2.0 / 3 * a.tC);   // :235
}

hipError_t launch_eij(const EijArgs& args, hipStream_t s) {
    const int64_t n = args.idx ? args.nidx : args.P;
    if (n == 0) return hipSuccess;
    EijArgs a = args;
    eij_params(a);
    const int64_t nblk = (n + 127) / 128;
    if (a.idx && a.nidx_dev)   // a device-side count: at most 2048 workgroups
        hipLaunchKernelGGL(eij_list_kernel, dim3((unsigned)std::min<int64_t>(nblk, 2048)), dim3(128), 0, s, a);
    else
        hipLaunchKernelGGL(eij_kernel, dim3((unsigned)nblk), dim3(128), 0, s, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (a.wave_flag && !a.idx)
        hipLaunchKernelGGL(eij_regularized_waves_kernel, dim3((unsigned)(((a.P + 63) / 64 + 255) / 256)), dim3(256),
                           0, s, a);
    else
        hipLaunchKernelGGL(eij_regularized_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 2048)),
                           dim3(256), 0, s, a);
    return hipGetLastError();
}

static void sachs3_params(EijArgs& a) {
    // rheo_params_tranisotropic (src/rheologies.f90:123-132) of the n' = 3 Sachs branch
    const double ne = 2.0 / (3.0 + 1);
    const double pc = std::pow(a.Ecc, ne), pa = std::pow(a.Eca, ne);
    a.sA = (pc - 1) / 2;
    a.sB = (3 * (pc + 1) - 2) / 2 - 2 * pa;
    a.sC = pa - 1;
}

hipError_t launch_eij_sachs3(const EijArgs& args, const Sachs3Args& b, hipStream_t s) {
    if (args.P == 0) return hipSuccess;
    EijArgs a = args;
    sachs3_params(a);
    hipLaunchKernelGGL(eij_sachs3_kernel<false>, dim3((unsigned)((a.P + 63) / 64)), dim3(64), 0, s, a, b);
    return hipGetLastError();
}

hipError_t launch_evw_sachs3(const EijArgs& args, const Sachs3Args& b, hipStream_t s) {
    if (args.P == 0 || args.nstates == 0) return hipSuccess;
    EijArgs a = args;
    sachs3_params(a);
    hipLaunchKernelGGL(eij_sachs3_kernel<true>, dim3((unsigned)((a.P + 63) / 64)), dim3(64), 0, s, a, b);
    return hipGetLastError();
}

hipError_t launch_evw(const EijArgs& args, hipStream_t s) {
    if (args.P == 0 || args.nstates == 0) return hipSuccess;
    EijArgs a = args;
    eij_params(a);
    hipLaunchKernelGGL(evw_kernel, dim3((unsigned)((a.P + 127) / 128)), dim3(128), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_ev_D4(int64_t P, int nrows, const double* A4F, const double* x, const double* tau, double* out,
                        hipStream_t s) {
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(ev_D4_kernel, dim3((unsigned)((P + 63) / 64)), dim3(64), 0, s, P, nrows, A4F, x, tau, out);
    return hipGetLastError();
}

hipError_t launch_moments68(int64_t P, int nrows, const double* x, double* m6, double* m8, hipStream_t s) {
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(moments68_kernel, dim3((unsigned)((P + 127) / 128)), dim3(128), 0, s, P, nrows, x, m6, m8);
    return hipGetLastError();
}

hipError_t launch_eig3(int64_t P, const double* M, double* ei, double* lami, hipStream_t s) {
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(eig3_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, P, M, ei, lami);
    return hipGetLastError();
}

hipError_t launch_caffe(int64_t P, int n_rss, double Emin, double Emax, double* DE, hipStream_t s) {
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(caffe_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, P, n_rss, Emin, Emax, DE);
    return hipGetLastError();
}

}  // namespace sfg
