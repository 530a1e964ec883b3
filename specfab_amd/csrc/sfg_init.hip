// Initial states from structure tensors, the DDRX rate factor and the
// power-spectrum bounds, batched over parcels (FP64).
//
//   a2_to_state_kernel  a2_to_nlm (src/moments.f90:68-74, body
//                       src/include/a2_to_nlm__body.f90): the l <= 2 rows of
//                       the real SoA state from a2; higher rows zeroed
//   a4_to_state_kernel  a4_to_nlm (src/moments.f90:76-84, body
//                       src/include/a4_to_nlm__body.f90): the l <= 4 rows from
//                       a4 (through its Mandel matrix, src/mandel.f90:52-66)
//   gamma0_kernel       Gamma0(eps, T, A, Q) (src/dynamics.f90:451-472)
//   bounds_kernel       apply_bounds (src/dynamics.f90:530-557) with Sl
//                       (src/idealstate.f90:80-93), in place on the state
//
// The reference evaluates these bodies with default-kind (REAL(4)) literals:
// Sqrt((5.0)), Sqrt(7.5), Sqrt((30.0)), Sqrt(2.5), Sqrt(17.5) are float32
// square roots, and products of two such literals (3.0 x Sqrt(5.0),
// Sqrt(2.5) x 3.0, Sqrt(17.5) x (+-3.0)) are rounded to float32 before they
// meet a REAL(8) operand; the gas constant 8.31446261815324 (no kind suffix)
// is its float32 value.  Everything else is FP64 in the reference's order of
// operations.  Real state layout per even l: [Re n_l0, Re n_l1, Im n_l1, ...,
// Re n_ll, Im n_ll].
#include <hip/hip_runtime.h>

#include "sfg_init.h"

namespace sfg {

namespace {

unsigned nb(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

constexpr double PI = 3.141592653589793;

// a REAL(4) square root widened to REAL(8)
__device__ inline double f32sqrt(float v) { return (double)sqrtf(v); }

}  // namespace

__global__ void a2_to_state_kernel(int64_t P, int n, const double* __restrict__ a2, double* __restrict__ x) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const double* A = a2 + p * 9;
    const double a11 = A[0], a22 = A[4], a12 = A[1], a13 = A[2], a23 = A[5];
    const double s5 = f32sqrt(5.0f), s75 = f32sqrt(7.5f), s30 = f32sqrt(30.0f);
    const double c3s5 = (double)(3.0f * sqrtf(5.0f));      // (3.0)*Sqrt((5.0)) in REAL(4)
    const double hs75 = (double)(0.5f * sqrtf(7.5f));       // (0,0.5)*Sqrt(7.5) in COMPLEX(4)
    const double sp = sqrt(PI), ip = sqrt(1.0 / PI);
    double v[6];
    v[0] = 1.0 / (2.0 * sp);
    v[1] = -((0.25 * ((-2.0 * s5 + c3s5 * a11) + c3s5 * a22)) / sp);
    v[2] = -(0.5 * ((s75 * a13) * ip));
    v[3] = (hs75 * a23) * ip;
    v[4] = -((0.125 * ((-s30) * a11 + s30 * a22)) / sp);
    v[5] = -((hs75 * a12) * ip);
    for (int q = 0; q < n; q++) x[(int64_t)q * P + p] = q < 6 ? v[q] : 0.0;
}

__global__ void a4_to_state_kernel(int64_t P, int n, const double* __restrict__ a4, double* __restrict__ x) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const double* A = a4 + p * 81;
    auto T = [&](int i, int j, int k, int l) { return A[((i * 3 + j) * 3 + k) * 3 + l]; };
    // Mandel 6x6 M[r][c] (0-based) of src/mandel.f90:60-65
    const double s = sqrt(2.0);
    const int pi_[6] = {0, 1, 2, 1, 0, 0}, pj_[6] = {0, 1, 2, 2, 2, 1};
    double M[6][6];
    for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++) {
            const double f = (r < 3 && c < 3) ? 1.0 : ((r < 3 || c < 3) ? s : 2.0);
            M[r][c] = (f == 1.0) ? T(pi_[r], pj_[r], pi_[c], pj_[c]) : f * T(pi_[r], pj_[r], pi_[c], pj_[c]);
        }
    const double sp = sqrt(PI), tsp = 2.0 * sp, ipi = 1.0 / PI, ip = sqrt(ipi);
    const double r5 = sqrt(5.0 * ipi), r10 = sqrt(10.0 * ipi), r15 = sqrt(15.0 * ipi), r35 = sqrt(35.0 * ipi),
                 r70 = sqrt(70.0 * ipi);
    const double s75 = f32sqrt(7.5f), s25 = f32sqrt(2.5f), s175 = f32sqrt(17.5f);
    const double hs15 = (double)(0.5f * sqrtf(15.0f));      // (0,0.5)*Sqrt((15.0))
    const double s25x3 = (double)(sqrtf(2.5f) * 3.0f);      // Sqrt(2.5)*(3.0)
    const double s175x3 = (double)(sqrtf(17.5f) * 3.0f);    // Sqrt(17.5)*(3.0)
    const double s175xm3 = (double)(sqrtf(17.5f) * -3.0f);  // Sqrt(17.5)*(-3.0)
    double v[15];
    v[0] = 1.0 / tsp;
    // l = 2: n20, n21, n22
    {
        const double t = ((((r5 - (3.0 * M[0][0] * r5) / 2.0) + (-3.0 * M[0][1]) * r5) - (3.0 * M[0][2] * r5) / 2.0) -
                          (3.0 * M[1][1] * r5) / 2.0) - (3.0 * M[1][2] * r5) / 2.0;
        v[1] = t / 2.0;
        const double im = hs15 * ((M[0][3] / tsp + M[1][3] / tsp) + M[2][3] / tsp);
        const double re = ((-(0.5 * (M[0][4] * r15)) - (M[1][4] * r15) / 2.0) - (M[2][4] * r15) / 2.0) / 2.0;
        v[2] = re;
        v[3] = im;
        const double re6 = ((((s75 * M[0][0]) * ip + (s75 * M[0][2]) * ip) + ((-s75) * M[1][1]) * ip) +
                            ((-s75) * M[1][2]) * ip) / 4.0;
        const double z6 = (-(0.5 * (M[0][5] * r15)) - (M[1][5] * r15) / 2.0) - (M[2][5] * r15) / 2.0;
        v[4] = re6;
        v[5] = 0.5 * z6;
    }
    // l = 4: n40 .. n44
    {
        const double w = (((((-4.0) / sp + (5.0 * M[0][0]) / tsp) + (5.0 * M[0][1]) / sp) + (20.0 * M[0][2]) / sp) +
                          (5.0 * M[1][1]) / tsp) + (20.0 * M[1][2]) / sp;
        v[6] = (-3.0 * w) / 8.0;
        const double v1 = ((s25x3 * M[0][3]) * ip + (s25x3 * M[1][3]) * ip) + (-2.0 * M[2][3]) * r10;
        const double v2 = ((s25x3 * M[0][4]) * ip + (s25x3 * M[1][4]) * ip) + (-2.0 * M[2][4]) * r10;
        v[7] = (3.0 * v2) / 8.0;
        v[8] = -0.375 * v1;
        const double u = (((s25 * M[0][0]) * ip + ((-s25) * M[1][1]) * ip) + (-3.0 * M[0][2]) * r10) +
                         (3.0 * M[1][2]) * r10;
        v[9] = (-3.0 * u) / 8.0;
        v[10] = 0.75 * (((M[0][5] * r5) / 2.0 + (M[1][5] * r5) / 2.0) + (-3.0 * M[2][5]) * r5);
        const double w1 = (s175x3 * M[0][3]) * ip + ((-s175) * M[1][3]) * ip;
        const double w2 = (s175 * M[0][4]) * ip + (s175xm3 * M[1][4]) * ip;
        v[11] = -((3.0 * w2) / 8.0);
        v[12] = 0.375 * w1;
        const double y1 = (M[0][5] * r35) / 2.0 - (M[1][5] * r35) / 2.0;
        const double y2 = ((s175 * M[0][0]) * ip + (s175 * M[1][1]) * ip) + (-3.0 * M[0][1]) * r70;
        v[13] = (3.0 * y2) / 16.0;
        v[14] = -0.75 * y1;
    }
    for (int q = 0; q < n; q++) x[(int64_t)q * P + p] = q < 15 ? v[q] : 0.0;
}

__global__ void gamma0_kernel(int64_t P, const double* __restrict__ eps, const double* __restrict__ T, double Tconst,
                              double A, double Q, double* __restrict__ G) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const double* e = eps + p * 9;
    // tr(eps . eps) as the diagonal of matmul(eps, eps), summed 1, 2, 3
    double d[3];
    for (int i = 0; i < 3; i++) {
        double s = 0.0;
        for (int k = 0; k < 3; k++) s += e[i * 3 + k] * e[k * 3 + i];
        d[i] = s;
    }
    const double I2 = sqrt(0.5 * ((d[0] + d[1]) + d[2]));
    const double R = (double)8.31446261815324f;
    const double Tp = T ? T[p] : Tconst;
    G[p] = A * I2 * exp(-Q / (R * Tp));
}

// S(l) of the real state block starting at row b: (1 / (2l + 1)) sum_m |n_lm|^2
__device__ inline double block_power(const double* x, int64_t P, int64_t p, int b, int l) {
    double s = x[(int64_t)b * P + p] * x[(int64_t)b * P + p];
    double t = 0.0;
    for (int m = 1; m <= l; m++) {
        const double re = x[(int64_t)(b + 2 * m - 1) * P + p], im = x[(int64_t)(b + 2 * m) * P + p];
        t += re * re + im * im;
    }
    return (s + 2.0 * t) / (double)(2 * l + 1);
}

__global__ void bounds_kernel(int64_t P, int n, double* __restrict__ x) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const double x0 = x[p];
    const double S0 = x0 * x0;
    for (int l = 2; l <= 4 && (l + 1) * (l + 2) / 2 <= n; l += 2) {
        const int b = (l == 2) ? 1 : 6;
        const double rel = block_power(x, P, p, b, l) / S0;
        if (rel > 1.0) {
            const double d = sqrt(rel);
            for (int q = b; q < b + 2 * l + 1; q++) x[(int64_t)q * P + p] /= d;
        }
    }
}

hipError_t launch_a2_to_state(int64_t P, int n, const double* a2, double* x, hipStream_t s) {
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(a2_to_state_kernel, dim3(nb(P, 128)), dim3(128), 0, s, P, n, a2, x);
    return hipGetLastError();
}

hipError_t launch_a4_to_state(int64_t P, int n, const double* a4, double* x, hipStream_t s) {
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(a4_to_state_kernel, dim3(nb(P, 128)), dim3(128), 0, s, P, n, a4, x);
    return hipGetLastError();
}

hipError_t launch_gamma0(int64_t P, const double* eps, const double* T, double Tconst, double A, double Q, double* G,
                         hipStream_t s) {
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(gamma0_kernel, dim3(nb(P, 256)), dim3(256), 0, s, P, eps, T, Tconst, A, Q, G);
    return hipGetLastError();
}

hipError_t launch_bounds(int64_t P, int n, double* x, hipStream_t s) {
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(bounds_kernel, dim3(nb(P, 256)), dim3(256), 0, s, P, n, x);
    return hipGetLastError();
}

}  // namespace sfg
