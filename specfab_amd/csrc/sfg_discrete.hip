// Discrete-grain path (FP64).
//
//   ri_lrot_kernel     ri_LROT (src/dynamics.f90:112-137): Euler steps of
//                      every grain's axis under the lattice-rotation
//                      velocity (W + iota (r r^T D - D r r^T)) r, each step
//                      renormalised; one lane per grain, the step's D and W
//                      uniform over the launch (scalar loads), the whole
//                      trajectory written (Nt, N, 3) -- HBM-write bound
//                      (24 bytes per grain-step).
//   ri_moments_*       ri_to_nlm (src/moments.f90:98-131): a group of 1 to 64
//                      lanes (ri_moments_group_kernel) or a whole workgroup
//                      (ri_moments_block_kernel) per ensemble sums the weighted
//                      degree-k monomials of its grains' axes (k = 2, 4 or
//                      6: the a2, a4 or a6 tensor, whose distinct entries
//                      they are), reduced across the group; k = 2, 4
//                      hand the tensor to a2_to_nlm /
//                      a4_to_nlm (sfg_init.hip), k = 6 applies the a6 map
//                      (a6_to_nlm, src/moments.f90:86-92, derived in
//                      tables/make_a6_map.py) and writes the l <= 6 rows of
//                      the real state.  n00 is 1 / (2 sqrt(pi)) whatever the
//                      weights, as the reference's bodies set it.
#include "sfg_discrete.h"

#include "generated/sfg_a6_map.h"

namespace sfg {

typedef double dbl2 __attribute__((ext_vector_type(2)));

// One lane per grain, 64 grains per wave.  Per step the velocity
// (W + iota (r r^T D - D r r^T)) r is formed as W r + iota (r (r.Dr) - Dr (r.r))
// (the same products, without the 3x3 intermediates), and the renormalisation
// multiplies by 1/|r_new| from the hardware reciprocal square root refined by
// two Newton steps y <- y + y (1 - s y^2) / 2 (within 1 ulp; the reference
// divides by norm2).  The pointers are restrict-qualified so that the step's
// D and W (uniform over the launch) come through scalar loads.  The wave's 64
// axes (1536 contiguous bytes of the trajectory) go out through LDS as
// 16-byte non-temporal stores: the trajectory is written once and not read
// back by this kernel.  A wave whose 64 axes do not start 16-byte aligned (N
// odd and t odd, or an unaligned output) or that holds fewer than 64 grains
// stores per lane.
__global__ void __launch_bounds__(256) ri_lrot_kernel(int64_t N, int Nt, double dt, double iota,
                                                      const double* __restrict__ ri0,
                                                      const double* __restrict__ Dm,
                                                      const double* __restrict__ Wm, double* __restrict__ ri) {
    __shared__ double stage[4][192];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t base = ((int64_t)blockIdx.x * 4 + w) * 64;    // first grain of the wave
    if (base >= N) return;
    const bool full = base + 64 <= N;
    const int64_t i = base + lane;
    const bool live = i < N;
    double r0 = 0.0, r1 = 0.0, r2 = 1.0;
    if (live) r0 = ri0[i * 3 + 0], r1 = ri0[i * 3 + 1], r2 = ri0[i * 3 + 2];
    double* st = stage[w];
    for (int t = 0; t < Nt; t++) {
        if (t > 0) {
            const double* D = Dm + (int64_t)(t - 1) * 9;
            const double* W = Wm + (int64_t)(t - 1) * 9;
            const double d0 = D[0] * r0 + D[1] * r1 + D[2] * r2;
            const double d1 = D[3] * r0 + D[4] * r1 + D[5] * r2;
            const double d2 = D[6] * r0 + D[7] * r1 + D[8] * r2;
            const double rDr = r0 * d0 + r1 * d1 + r2 * d2;
            const double rr = r0 * r0 + r1 * r1 + r2 * r2;
            const double v0 = (W[0] * r0 + W[1] * r1 + W[2] * r2) + iota * (r0 * rDr - d0 * rr);
            const double v1 = (W[3] * r0 + W[4] * r1 + W[5] * r2) + iota * (r1 * rDr - d1 * rr);
            const double v2 = (W[6] * r0 + W[7] * r1 + W[8] * r2) + iota * (r2 * rDr - d2 * rr);
            const double n0 = r0 + dt * v0, n1 = r1 + dt * v1, n2 = r2 + dt * v2;
            const double ss = n0 * n0 + n1 * n1 + n2 * n2;
            double y = __builtin_amdgcn_rsq(ss);
            double e = fma(-ss * y, y, 1.0);
            y = fma(0.5 * y, e, y);
            e = fma(-ss * y, y, 1.0);
            y = fma(0.5 * y, e, y);
            r0 = n0 * y, r1 = n1 * y, r2 = n2 * y;
        }
        double* row = ri + (int64_t)t * N * 3;
        const int64_t off = base * 3;                            // doubles from the row start
        if (full && ((uintptr_t)(row + off) & 15) == 0) {
            st[lane * 3 + 0] = r0, st[lane * 3 + 1] = r1, st[lane * 3 + 2] = r2;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            dbl2* dst = (dbl2*)(row + off);
            const dbl2 lo = {st[lane * 2], st[lane * 2 + 1]};
            __builtin_nontemporal_store(lo, dst + lane);
            if (lane < 32) {
                const dbl2 hi = {st[128 + lane * 2], st[128 + lane * 2 + 1]};
                __builtin_nontemporal_store(hi, dst + 64 + lane);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else if (live) {
            row[i * 3 + 0] = r0, row[i * 3 + 1] = r1, row[i * 3 + 2] = r2;
        }
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

// the real state rows (k = 6) or the tensor (k = 2, 4) of ensemble e from its
// summed monomials mom[]: entries first, first + stride, ...; with LANE (mom
// in registers) one thread writes every entry through fully unrolled loops,
// so that mom is indexed by constants only
template <int K, bool LANE>
__device__ __forceinline__ void moments_out(const RiMomentArgs& A, int64_t e, const double* mom, int first,
                                            int stride) {
    if (LANE) first = 0, stride = 1;
    if (K == 2) {
#pragma unroll
        for (int c = first; c < 9; c += stride) {
            const int i = c / 3, j = c % 3;
            int ex[3] = {0, 0, 0};
            ex[i]++, ex[j]++;
            A.out[e * 9 + c] = mom[mono_index<K>(ex[0], ex[1])];
        }
    } else if (K == 4) {
#pragma unroll
        for (int c = first; c < 81; c += stride) {
            int t = c, ex[3] = {0, 0, 0};
#pragma unroll
            for (int d = 0; d < 4; d++) {
                ex[t % 3]++;
                t /= 3;
            }
            A.out[e * 81 + c] = mom[mono_index<K>(ex[0], ex[1])];
        }
    } else {
        constexpr double PI = 3.141592653589793;
        const int top = A.n < 28 ? A.n : 28;
#pragma unroll
        for (int q = first; q < 28; q += stride) {
            if (q >= top) break;
            double v;
            if (q == 0) {
                v = 1.0 / (2.0 * sqrt(PI));
            } else {
                const double ipi = 1.0 / PI;
                double sum = 0.0;
#pragma unroll
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
        for (int q = 28 + first; q < A.n; q += stride) A.out[(int64_t)q * A.E + e] = 0.0;
    }
}

// weighted monomial sums of grains g0, g0 + gs, ... < N of ensemble R into s[],
// four grains' loads in flight at a time (a grain past N reloads grain g < N
// and counts with weight 0)
template <int K>
__device__ __forceinline__ void monomial_sums(const double* __restrict__ R, const double* __restrict__ wi, int64_t N,
                                              int64_t g0, int gs, double* s) {
    constexpr int U = 4;
    for (int64_t g = g0; g < N; g += (int64_t)U * gs) {
        double x[U], y[U], z[U], w[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int64_t gg = g + (int64_t)u * gs;
            const bool ok = gg < N;
            const int64_t gc = ok ? gg : g;
            x[u] = R[gc * 3 + 0], y[u] = R[gc * 3 + 1], z[u] = R[gc * 3 + 2];
            w[u] = ok ? wi[gc] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            double px[K + 1], py[K + 1], pz[K + 1];
            px[0] = py[0] = pz[0] = 1.0;
#pragma unroll
            for (int q = 1; q <= K; q++) {
                px[q] = px[q - 1] * x[u];
                py[q] = py[q - 1] * y[u];
                pz[q] = pz[q - 1] * z[u];
            }
            int k = 0;
#pragma unroll
            for (int aa = K; aa >= 0; aa--)
#pragma unroll
                for (int bb = K - aa; bb >= 0; bb--) {
                    s[k] += w[u] * (px[aa] * py[bb] * pz[K - aa - bb]);
                    k++;
                }
        }
    }
}

// TPE lanes per ensemble (TPE = 1, 2, ..., 64; 256 / TPE ensembles per
// workgroup): each lane sums a stride-TPE subset of the grains, the group
// all-reduces with lane shuffles, and the group's first lane writes the
// outputs (mom in registers, indexed by constants only)
template <int K, int TPE>
__device__ void ri_moments_group(const RiMomentArgs& A) {
    constexpr int NM = Mono<K>::count;
    const int64_t e = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / TPE;
    const int sub = threadIdx.x % TPE;
    double s[NM];
#pragma unroll
    for (int k = 0; k < NM; k++) s[k] = 0.0;
    if (e < A.E) monomial_sums<K>(A.ri + e * A.N * 3, A.wi, A.N, sub, TPE, s);
#pragma unroll
    for (int k = 0; k < NM; k++)
#pragma unroll
        for (int off = TPE / 2; off > 0; off >>= 1) s[k] += __shfl_xor(s[k], off, 64);
    if (e < A.E && sub == 0) moments_out<K, true>(A, e, s, 0, 1);
}

// one 256-thread workgroup per ensemble (few, large ensembles): every
// thread's sums go to LDS, 8 threads per monomial add 32 of them each, then
// one thread per monomial adds the 8 partial sums
template <int K>
__device__ void ri_moments_block(const RiMomentArgs& A) {
    constexpr int NM = Mono<K>::count;
    __shared__ double buf[NM][257];
    __shared__ double part[NM][8];
    __shared__ double mom[NM];
    const int64_t e = blockIdx.x;
    double s[NM];
#pragma unroll
    for (int k = 0; k < NM; k++) s[k] = 0.0;
    monomial_sums<K>(A.ri + e * A.N * 3, A.wi, A.N, threadIdx.x, 256, s);
#pragma unroll
    for (int k = 0; k < NM; k++) buf[k][threadIdx.x] = s[k];
    __syncthreads();
    if (threadIdx.x < NM * 8) {
        const int k = threadIdx.x / 8, c = threadIdx.x % 8;
        double v = 0.0;
        for (int j = 0; j < 32; j++) v += buf[k][c * 32 + j];
        part[k][c] = v;
    }
    __syncthreads();
    if (threadIdx.x < NM) {
        double v = 0.0;
        for (int c = 0; c < 8; c++) v += part[threadIdx.x][c];
        mom[threadIdx.x] = v;
    }
    __syncthreads();
    moments_out<K, false>(A, e, mom, threadIdx.x, 256);
}

template <int TPE>
__global__ void __launch_bounds__(256) ri_moments_group_kernel(RiMomentArgs a) {
    if (a.order == 2) ri_moments_group<2, TPE>(a);
    else if (a.order == 4) ri_moments_group<4, TPE>(a);
    else ri_moments_group<6, TPE>(a);
}

// one kernel per order, so that each holds only its own LDS buffers
template <int K>
__global__ void __launch_bounds__(256) ri_moments_block_kernel(RiMomentArgs a) {
    ri_moments_block<K>(a);
}

}  // namespace

#ifndef SFG_HOST_EMUL
hipError_t launch_ri_lrot(const RiLrotArgs& a, hipStream_t s) {
    if (a.N == 0) return hipSuccess;
    hipLaunchKernelGGL(ri_lrot_kernel, dim3((unsigned)((a.N + 255) / 256)), dim3(256), 0, s,   // 4 waves of 64 grains
                       a.N, a.Nt, a.dt, a.iota, a.ri0, a.D, a.W, a.ri);
    return hipGetLastError();
}

hipError_t launch_ri_moments(const RiMomentArgs& a, hipStream_t s) {
    if (a.E == 0) return hipSuccess;
    // lanes per ensemble: enough for about 2^18 threads in flight, at most 64
    // and at most N / 16 (each lane keeps >= 16 grains); a whole workgroup per
    // ensemble when that asks for >= 256 lanes and each keeps >= 4 grains
    const int64_t target = ((int64_t)1 << 18) / a.E;
    if (target >= 256 && a.N >= 1024) {
        const dim3 grid((unsigned)a.E), block(256);
        if (a.order == 2) hipLaunchKernelGGL(ri_moments_block_kernel<2>, grid, block, 0, s, a);
        else if (a.order == 4) hipLaunchKernelGGL(ri_moments_block_kernel<4>, grid, block, 0, s, a);
        else hipLaunchKernelGGL(ri_moments_block_kernel<6>, grid, block, 0, s, a);
        return hipGetLastError();
    }
    int tpe = 1;
    while (tpe < 64 && 2 * tpe <= target && 2 * tpe * 16 <= a.N) tpe *= 2;
    const dim3 grid((unsigned)((a.E * tpe + 255) / 256)), block(256);
    switch (tpe) {
        case 1: hipLaunchKernelGGL(ri_moments_group_kernel<1>, grid, block, 0, s, a); break;
        case 2: hipLaunchKernelGGL(ri_moments_group_kernel<2>, grid, block, 0, s, a); break;
        case 4: hipLaunchKernelGGL(ri_moments_group_kernel<4>, grid, block, 0, s, a); break;
        case 8: hipLaunchKernelGGL(ri_moments_group_kernel<8>, grid, block, 0, s, a); break;
        case 16: hipLaunchKernelGGL(ri_moments_group_kernel<16>, grid, block, 0, s, a); break;
        case 32: hipLaunchKernelGGL(ri_moments_group_kernel<32>, grid, block, 0, s, a); break;
        default: hipLaunchKernelGGL(ri_moments_group_kernel<64>, grid, block, 0, s, a); break;
    }
    return hipGetLastError();
}
#endif

}  // namespace sfg
