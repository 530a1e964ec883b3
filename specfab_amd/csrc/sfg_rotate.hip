// Rotation of states by Wigner-D matrices (rotate_nlm, src/rotation.f90:29-55),
// eigentensors of the fourth-order structure tensor (a4_eigentensors,
// src/frames.f90:131-152) and state validity (nlm_isvalid,
// src/idealstate.f90:50-78), FP64, one lane per parcel.
//
// Rotation: n'_l^m = sum_n exp(-i n phi) d^l_{mn}(theta) n_l^n per degree l
// (the reference's D^l(m, n), src/include/Dlmn.f90).  The small-d entries
// come from Wigner's sum formula, d = sum_k coef_k cos(theta/2)^a sin(theta/2)^b,
// with the coefficients tabulated on the host (wigner_terms) and every loop
// over (l, m, n, k) unrolled at compile time, so that the exponents are
// constants and the coefficients scalar loads.  In the real basis only m >= 0
// is formed: with u_n + i v_n = n_l^n exp(-i n phi),
//   Re n'_m = d_m0 n_0 + sum_{n>0} (d_mn + (-1)^n d_m,-n) u_n,
//   Im n'_m =            sum_{n>0} (d_mn - (-1)^n d_m,-n) v_n.
#include <type_traits>

#include "sfg_rotate.h"

namespace sfg {

namespace {

#ifndef SFG_HOST_EMUL
unsigned nb(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
#endif

template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

template <int l, int m, int n, int K>
__device__ __forceinline__ double dsmall(const double* __restrict__ coef, const double (&cp)[K],
                                         const double (&sp)[K]) {
    constexpr int T = wigner_offset(l, m, n);
    constexpr int k0 = n - m > 0 ? n - m : 0;
    constexpr int k1 = (l + n < l - m) ? l + n : l - m;
    double acc = 0.0;
    static_for<k0, k1 + 1>([&](auto kk) {
        constexpr int k = kk;
        acc = fma(coef[T + k - k0], cp[2 * l + n - m - 2 * k] * sp[m - n + 2 * k], acc);
    });
    return acc;
}

template <int l, int K>
__device__ __forceinline__ void rotate_block(const RotateArgs& a, int64_t p, double phi, const double (&cp)[K],
                                             const double (&sp)[K]) {
    constexpr int base = (l - 1) * l / 2;   // first real-basis row of degree l
    const int64_t P = a.P;
    double u[l + 1], v[l + 1];
    const double x0 = a.x[(int64_t)base * P + p];
    static_for<1, l + 1>([&](auto nn) {
        constexpr int n = nn;
        double sn, cn;
        sincos((double)n * phi, &sn, &cn);
        const double re = a.x[(int64_t)(base + 2 * n - 1) * P + p], im = a.x[(int64_t)(base + 2 * n) * P + p];
        u[n] = cn * re + sn * im;
        v[n] = cn * im - sn * re;
    });
    // the whole block is read before any of it is written (out may be x)
    double o[2 * l + 1];
    static_for<0, l + 1>([&](auto mm) {
        constexpr int m = mm;
        double re = dsmall<l, m, 0>(a.coef, cp, sp) * x0, im = 0.0;
        static_for<1, l + 1>([&](auto nn) {
            constexpr int n = nn;
            const double dpos = dsmall<l, m, n>(a.coef, cp, sp), dneg = dsmall<l, m, -n>(a.coef, cp, sp);
            const double sg = (n & 1) ? -1.0 : 1.0;
            re = fma(dpos + sg * dneg, u[n], re);
            im = fma(dpos - sg * dneg, v[n], im);
        });
        if constexpr (m == 0) {
            o[0] = re;
        } else {
            o[2 * m - 1] = re;
            o[2 * m] = im;
        }
    });
    static_for<0, 2 * l + 1>([&](auto ii) { a.out[(int64_t)(base + (int)ii) * P + p] = o[(int)ii]; });
}

}  // namespace

template <int LM>
__global__ void __launch_bounds__(128) rotate_kernel(RotateArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.P) return;
    const double th = a.theta[p], ph = a.phi[p];
    const double c = cos(th / 2), s = sin(th / 2);
    constexpr int K = 2 * LM + 1;
    double cp[K], sp[K];
    cp[0] = 1.0;
    sp[0] = 1.0;
    static_for<1, K>([&](auto kk) {
        constexpr int k = kk;
        cp[k] = cp[k - 1] * c;
        sp[k] = sp[k - 1] * s;
    });
    a.out[p] = a.x[p];   // l = 0
    static_for<1, LM / 2 + 1>([&](auto bb) {
        constexpr int l = 2 * bb;
        rotate_block<l>(a, p, ph, cp, sp);
    });
}

#ifdef SFG_HOST_EMUL
template void rotate_kernel<2>(RotateArgs);
template void rotate_kernel<4>(RotateArgs);
template void rotate_kernel<6>(RotateArgs);
template void rotate_kernel<8>(RotateArgs);
template void rotate_kernel<10>(RotateArgs);
template void rotate_kernel<12>(RotateArgs);
#else
hipError_t launch_rotate(const RotateArgs& a, hipStream_t s) {
    if (a.P == 0) return hipSuccess;
    const dim3 g(nb(a.P, 128)), b(128);
    switch (a.L) {
        case 2: hipLaunchKernelGGL(rotate_kernel<2>, g, b, 0, s, a); break;
        case 4: hipLaunchKernelGGL(rotate_kernel<4>, g, b, 0, s, a); break;
        case 6: hipLaunchKernelGGL(rotate_kernel<6>, g, b, 0, s, a); break;
        case 8: hipLaunchKernelGGL(rotate_kernel<8>, g, b, 0, s, a); break;
        case 10: hipLaunchKernelGGL(rotate_kernel<10>, g, b, 0, s, a); break;
        case 12: hipLaunchKernelGGL(rotate_kernel<12>, g, b, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
#endif

// ------------------------------------------------------------ a4 eigentensors

namespace {

// entry (of the 15 index-sorted a4 entries 1111, 1112, 1113, 1122, 1123, 1133,
// 1222, 1223, 1233, 1333, 2222, 2223, 2233, 2333, 3333) at Mandel (I, J),
// pairs I = 11, 22, 33, 23, 13, 12 (a4_to_mat, src/mandel.f90:52-66)
__constant__ int kMandelEntry[6][6] = {{0, 3, 5, 4, 2, 1},    {3, 10, 12, 11, 7, 6}, {5, 12, 14, 13, 9, 8},
                                       {4, 11, 13, 12, 8, 7}, {2, 7, 9, 8, 5, 4},    {1, 6, 8, 7, 4, 3}};

// cyclic Jacobi of a symmetric N x N matrix; eigenvalues ascending (dsyev's
// order) with their eigenvectors as the columns of V.  Sweeps stop once the
// off-diagonal squares are below 1e-34 of the diagonal's.
template <int N>
__device__ void jacobi_sym(double (&A)[N][N], double (&V)[N][N], double (&lam)[N]) {
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) V[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 50; sweep++) {
        double off = 0.0, dia = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) {
            dia += A[i][i] * A[i][i];
#pragma unroll
            for (int j = i + 1; j < N; j++) off += A[i][j] * A[i][j];
        }
        if (!(off > 1e-34 * dia)) break;
#pragma unroll
        for (int p = 0; p < N - 1; p++)
#pragma unroll
            for (int q = p + 1; q < N; q++) {
                if (A[p][q] == 0.0) continue;
                const double th = (A[q][q] - A[p][p]) / (2 * A[p][q]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
                const double c = 1 / sqrt(t * t + 1), s = t * c;
#pragma unroll
                for (int k = 0; k < N; k++) {
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
#pragma unroll
                for (int k = 0; k < N; k++) {
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
#pragma unroll
                for (int k = 0; k < N; k++) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
#pragma unroll
    for (int i = 0; i < N; i++) lam[i] = A[i][i];
    // selection sort, ascending, carrying the eigenvector columns
#pragma unroll
    for (int i = 0; i < N - 1; i++) {
#pragma unroll
        for (int j = i + 1; j < N; j++) {
            const bool sw = lam[j] < lam[i];
            const double ti = lam[i];
            lam[i] = sw ? lam[j] : ti;
            lam[j] = sw ? ti : lam[j];
#pragma unroll
            for (int k = 0; k < N; k++) {
                const double vi = V[k][i];
                V[k][i] = sw ? V[k][j] : vi;
                V[k][j] = sw ? vi : V[k][j];
            }
        }
    }
}

}  // namespace

__global__ void __launch_bounds__(128) a4eig_kernel(A4EigArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.P) return;
    double xq[15];
#pragma unroll
    for (int q = 0; q < 15; q++) xq[q] = a.x[(int64_t)q * a.P + p];
    const double ix0 = 1.0 / xq[0];
    // the full a4 (REAL(4) body), then a4_to_mat with sqrt(2) and 2 weights
    double e4[15];
#pragma unroll
    for (int e = 0; e < 15; e++) {
        double s = 0;
#pragma unroll
        for (int q = 0; q < 15; q++) s += a.A4F[e * 15 + q] * xq[q];
        e4[e] = s * ix0;
    }
    const double r2 = sqrt(2.0);
    double M[6][6], V[6][6], lam[6];
#pragma unroll
    for (int I = 0; I < 6; I++)
#pragma unroll
        for (int J = 0; J < 6; J++) {
            const double w = (I < 3 ? 1.0 : r2) * (J < 3 ? 1.0 : r2);
            M[I][J] = (I >= 3 && J >= 3 ? 2.0 : w) * e4[kMandelEntry[I][J]];
        }
    jacobi_sym<6>(M, V, lam);
    if (a.ev)
#pragma unroll
        for (int i = 0; i < 6; i++) a.ev[p * 6 + i] = lam[i];
    if (a.Q)
#pragma unroll
        for (int i = 0; i < 6; i++) {
            // vec_to_mat (src/mandel.f90:26-37) of eigenvector i
            double* Q = a.Q + (p * 6 + i) * 9;
            const double v0 = V[0][i], v1 = V[1][i], v2 = V[2][i];
            const double v3 = V[3][i] / r2, v4 = V[4][i] / r2, v5 = V[5][i] / r2;
            Q[0] = v0; Q[1] = v5; Q[2] = v4;
            Q[3] = v5; Q[4] = v1; Q[5] = v3;
            Q[6] = v4; Q[7] = v3; Q[8] = v2;
        }
    if (a.valid) {
        // a2 eigenvalues (src/idealstate.f90:69-70) and a4 eigenvalues (:72-73) inside (0, 1)
        double a2v[6];
#pragma unroll
        for (int i = 0; i < 6; i++) {
            double s = 0;
#pragma unroll
            for (int q = 0; q < 6; q++) s += a.A2[i * 15 + q] * xq[q];
            a2v[i] = s * ix0 + (i < 3 ? 1.0 / 3.0 : 0.0);
        }
        double A3[3][3] = {{a2v[0], a2v[5] / r2, a2v[4] / r2},
                           {a2v[5] / r2, a2v[1], a2v[3] / r2},
                           {a2v[4] / r2, a2v[3] / r2, a2v[2]}};
        double V3[3][3], l3[3];
        jacobi_sym<3>(A3, V3, l3);
        const bool ok2 = l3[2] < 1.0 && l3[0] > 0.0;
        const bool ok4 = lam[5] < 1.0 && lam[0] > 0.0;
        a.valid[p] = (ok2 && ok4) ? 1 : 0;
    }
}

#ifndef SFG_HOST_EMUL
hipError_t launch_a4eig(const A4EigArgs& a, hipStream_t s) {
    if (a.P == 0) return hipSuccess;
    hipLaunchKernelGGL(a4eig_kernel, dim3(nb(a.P, 128)), dim3(128), 0, s, a);
    return hipGetLastError();
}
#endif

}  // namespace sfg
