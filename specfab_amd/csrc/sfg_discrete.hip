// Discrete-grain path (FP64).
//
//   ri_lrot_kernel     ri_LROT (src/dynamics.f90:112-137): Euler steps of
//                      every grain's axis under the lattice-rotation
//                      velocity (W + iota (r r^T D - D r r^T)) r, each step
//                      renormalised; one lane per grain, the step's D and W
//                      uniform over the launch (scalar loads), the whole
//                      trajectory written (Nt, N, 3).
//   ri_moments_kernel  ri_to_nlm (src/moments.f90:98-131): one workgroup per
//                      ensemble sums the weighted degree-k monomials of its
//                      grains' axes (k = 2, 4 or 6: the a2, a4 or a6 tensor,
//                      whose distinct entries they are), reduced across the
//                      workgroup; k = 2, 4 hand the tensor to a2_to_nlm /
//                      a4_to_nlm (sfg_init.hip), k = 6 applies the a6 map
//                      (a6_to_nlm, src/moments.f90:86-92, derived in
//                      tables/make_a6_map.py) and writes the l <= 6 rows of
//                      the real state.  n00 is 1 / (2 sqrt(pi)) whatever the
//                      weights, as the reference's bodies set it.
#include "sfg_discrete.h"

#include "generated/sfg_a6_map.h"

namespace sfg {

__global__ void __launch_bounds__(256) ri_lrot_kernel(RiLrotArgs a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    double r[3] = {a.ri0[i * 3 + 0], a.ri0[i * 3 + 1], a.ri0[i * 3 + 2]};
    for (int c = 0; c < 3; c++) a.ri[i * 3 + c] = r[c];
    for (int t = 0; t + 1 < a.Nt; t++) {
        const double* D = a.D + (int64_t)t * 9;
        const double* W = a.W + (int64_t)t * 9;
        // mm = r r^T; Wp = iota (mm D - D mm); velocity (W + Wp) r
        double mm[3][3], G[3][3];
        for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++) mm[p][q] = r[p] * r[q];
        for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++) {
                double md = 0.0, dm = 0.0;
                for (int k = 0; k < 3; k++) {
                    md += mm[p][k] * D[k * 3 + q];
                    dm += D[p * 3 + k] * mm[k][q];
                }
                G[p][q] = W[p * 3 + q] + a.iota * (md - dm);
            }
        double rn[3];
        for (int p = 0; p < 3; p++) {
            double v = 0.0;
            for (int k = 0; k < 3; k++) v += G[p][k] * r[k];
            rn[p] = r[p] + a.dt * v;
        }
        const double nrm = sqrt(rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2]);
        for (int p = 0; p < 3; p++) r[p] = rn[p] / nrm;
        double* o = a.ri + ((int64_t)(t + 1) * a.N + i) * 3;
        for (int c = 0; c < 3; c++) o[c] = r[c];
    }
}

namespace {

// number of monomials x^a y^b z^c with a + b + c = K, and their order:
// a from K down to 0, then b from K - a down to 0 (the order of M6)
template <int K>
struct Mono {
    static constexpr int count = (K + 1) * (K + 2) / 2;
};

template <int K>
__device__ __forceinline__ int mono_index(int a, int b) {
    int idx = 0;
    for (int ap = K; ap > a; ap--) idx += K - ap + 1;
    return idx + (K - a - b);
}

template <int K>
__device__ void ri_moments(const RiMomentArgs& A) {
    constexpr int NM = Mono<K>::count;
    __shared__ double part[4][NM];
    const int64_t e = blockIdx.x;
    const double* R = A.ri + e * A.N * 3;
    double s[NM];
#pragma unroll
    for (int k = 0; k < NM; k++) s[k] = 0.0;
    for (int64_t g = threadIdx.x; g < A.N; g += blockDim.x) {
        const double x = R[g * 3 + 0], y = R[g * 3 + 1], z = R[g * 3 + 2], w = A.wi[g];
        double px[K + 1], py[K + 1], pz[K + 1];
        px[0] = py[0] = pz[0] = 1.0;
#pragma unroll
        for (int q = 1; q <= K; q++) {
            px[q] = px[q - 1] * x;
            py[q] = py[q - 1] * y;
            pz[q] = pz[q - 1] * z;
        }
        int k = 0;
#pragma unroll
        for (int aa = K; aa >= 0; aa--)
#pragma unroll
            for (int bb = K - aa; bb >= 0; bb--) {
                s[k] += w * (px[aa] * py[bb] * pz[K - aa - bb]);
                k++;
            }
    }
    // wavefront sums, then the four waves' through LDS
#pragma unroll
    for (int k = 0; k < NM; k++)
        for (int off = 32; off > 0; off >>= 1) s[k] += __shfl_xor(s[k], off, 64);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NM; k++) part[wave][k] = s[k];
    __syncthreads();
    __shared__ double mom[NM];
    if (threadIdx.x < NM) {
        const int k = threadIdx.x;
        mom[k] = ((part[0][k] + part[1][k]) + part[2][k]) + part[3][k];
    }
    __syncthreads();
    if (K == 2 && threadIdx.x < 9) {
        const int i = threadIdx.x / 3, j = threadIdx.x % 3;
        int ex[3] = {0, 0, 0};
        ex[i]++, ex[j]++;
        A.out[e * 9 + threadIdx.x] = mom[mono_index<K>(ex[0], ex[1])];
    } else if (K == 4 && threadIdx.x < 81) {
        int t = threadIdx.x, ex[3] = {0, 0, 0};
        for (int d = 0; d < 4; d++) {
            ex[t % 3]++;
            t /= 3;
        }
        A.out[e * 81 + threadIdx.x] = mom[mono_index<K>(ex[0], ex[1])];
    } else if (K == 6) {
        constexpr double PI = 3.141592653589793;
        for (int q = threadIdx.x; q < A.n; q += blockDim.x) {
            double v = 0.0;
            if (q == 0) {
                v = 1.0 / (2.0 * sqrt(PI));
            } else if (q < 28) {
                const double ipi = 1.0 / PI;
                double sum = 0.0;
                for (int k = 0; k < 28; k++) sum += W6[q][k] * mom[k];
                // the row's normalising root: FP64, or a REAL(4) literal root
                // (times 3 in REAL(4) for kind 3)
                const float r32 = sqrtf((float)A6_Q[q]);
                const double Kq = A6_KIND[q] == 0   ? 1.0 / sqrt(PI)
                                  : A6_KIND[q] == 1 ? sqrt(A6_Q[q] * ipi)
                                  : A6_KIND[q] == 2 ? (double)r32 * sqrt(ipi)
                                                    : (double)(3.0f * r32) / 3.0 * sqrt(ipi);
                v = Kq * sum;
            }
            A.out[(int64_t)q * A.E + e] = v;
        }
    }
}

}  // namespace

__global__ void __launch_bounds__(256) ri_moments_kernel(RiMomentArgs a) {
    if (a.order == 2) ri_moments<2>(a);
    else if (a.order == 4) ri_moments<4>(a);
    else ri_moments<6>(a);
}

#ifndef SFG_HOST_EMUL
hipError_t launch_ri_lrot(const RiLrotArgs& a, hipStream_t s) {
    if (a.N == 0) return hipSuccess;
    hipLaunchKernelGGL(ri_lrot_kernel, dim3((unsigned)((a.N + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_ri_moments(const RiMomentArgs& a, hipStream_t s) {
    if (a.E == 0) return hipSuccess;
    hipLaunchKernelGGL(ri_moments_kernel, dim3((unsigned)a.E), dim3(256), 0, s, a);
    return hipGetLastError();
}
#endif

}  // namespace sfg
