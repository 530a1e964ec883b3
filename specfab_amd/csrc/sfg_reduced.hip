// Reduced real form, batched over mesh nodes (FP64, index work only).
//
// A real-valued ODF has n_l^{-m} = (-1)^m conj(n_l^m), so its expansion is
// fixed by the m >= 0 coefficients rnlm = (n_0^0, n_2^0, n_2^1, n_2^2,
// n_4^0, ...) (src/reducedform.f90:37-44, 160-174).  Coefficient k of rnlm
// with h = floor(sqrt(k)) is (l, m) = (2h, k - h^2); (l, m) sits at position
// l(l-1)/2 + l + m of the full nlm (0-based).
//
//   reduce_M_kernel     reduce_M (src/reducedform.f90:76-120): the rows of M
//                       at the m >= 0 positions, and for column (l, m) its +m
//                       and -m columns Zp, Zn with s = (-1)^m (0 for m = 0):
//                       Mrr = Re Zp + s Re Zn, Mri = -Im Zp + s Im Zn,
//                       Mir = Im Zp + s Im Zn, Mii = Re Zp - s Re Zn, so that
//                       d(rnlm)/dt = Mrr Re r + Mri Im r + i (Mir Re r + Mii Im r)
//   nlm_to_rnlm_kernel  nlm_to_rnlm (src/reducedform.f90:176-183)
//   rnlm_to_nlm_kernel  rnlm_to_nlm (src/reducedform.f90:160-174): m >= 0
//                       copied, m < 0 as (-1)^m conj(n_l^{|m|})
//
// One thread per output entry; entries of one node are contiguous, so the
// stores coalesce and the gathers of a row of M stay within one l block.
#include <hip/hip_runtime.h>

#include "sfg_reduced.h"

namespace sfg {

namespace {

unsigned nb(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// (l, m >= 0) of reduced coefficient k
__device__ inline void rlm_of(int k, int& l, int& m) {
    int h = (int)sqrt((double)k);
    while ((h + 1) * (h + 1) <= k) ++h;
    while (h * h > k) --h;
    l = 2 * h;
    m = k - h * h;
}

__device__ inline int full_index(int l, int m) { return l * (l - 1) / 2 + l + m; }

__global__ void reduce_M_kernel(int64_t P, int n, int rn, const double* __restrict__ M, double* __restrict__ Mrr,
                                double* __restrict__ Mri, double* __restrict__ Mir, double* __restrict__ Mii) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t per = (int64_t)rn * rn;
    if (t >= P * per) return;
    int64_t p = t / per;
    int e = (int)(t - p * per);
    int i = e / rn, j = e - i * rn;
    int li, mi, lj, mj;
    rlm_of(i, li, mi);
    rlm_of(j, lj, mj);
    const double* row = M + 2 * ((size_t)p * n * n + (size_t)full_index(li, mi) * n);
    int cp = full_index(lj, mj), cn = cp - 2 * mj;
    double vp = row[2 * cp], qp = row[2 * cp + 1];
    double vn = row[2 * cn], qn = row[2 * cn + 1];
    double s = (mj == 0) ? 0.0 : ((mj & 1) ? -1.0 : 1.0);
    Mrr[t] = vp + s * vn;
    Mri[t] = -qp + s * qn;
    Mir[t] = qp + s * qn;
    Mii[t] = vp - s * vn;
}

__global__ void nlm_to_rnlm_kernel(int64_t P, int n, int rn, const double* __restrict__ nlm,
                                   double* __restrict__ rnlm) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P * rn) return;
    int64_t p = t / rn;
    int k = (int)(t - p * rn), l, m;
    rlm_of(k, l, m);
    const double* src = nlm + 2 * ((size_t)p * n + full_index(l, m));
    rnlm[2 * t] = src[0];
    rnlm[2 * t + 1] = src[1];
}

__global__ void rnlm_to_nlm_kernel(int64_t P, int n, int rn, const double* __restrict__ rnlm,
                                   double* __restrict__ nlm) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P * n) return;
    int64_t p = t / n;
    int j = (int)(t - p * n);
    int l = 0;
    while ((l + 2) * (l + 1) / 2 <= j) l += 2;  // block of l holds positions l(l-1)/2 .. (l+1)(l+2)/2 - 1
    int m = j - full_index(l, 0);
    int am = m < 0 ? -m : m;
    const double* src = rnlm + 2 * ((size_t)p * rn + l * l / 4 + am);
    double re = src[0], im = src[1];
    if (m < 0) {
        double sg = (am & 1) ? -1.0 : 1.0;
        re = sg * re;
        im = -sg * im;
    }
    nlm[2 * t] = re;
    nlm[2 * t + 1] = im;
}

}  // namespace

hipError_t launch_reduce_M(int64_t P, int L, const double* M, double* Mrr, double* Mri, double* Mir, double* Mii,
                           hipStream_t s) {
    if (P == 0) return hipSuccess;
    int n = (L + 1) * (L + 2) / 2, rn = rnlm_len(L);
    hipLaunchKernelGGL(reduce_M_kernel, dim3(nb(P * rn * rn, 256)), dim3(256), 0, s, P, n, rn, M, Mrr, Mri, Mir,
                       Mii);
    return hipGetLastError();
}

hipError_t launch_nlm_to_rnlm(int64_t P, int L, const double* nlm, double* rnlm, hipStream_t s) {
    if (P == 0) return hipSuccess;
    int n = (L + 1) * (L + 2) / 2, rn = rnlm_len(L);
    hipLaunchKernelGGL(nlm_to_rnlm_kernel, dim3(nb(P * rn, 256)), dim3(256), 0, s, P, n, rn, nlm, rnlm);
    return hipGetLastError();
}

hipError_t launch_rnlm_to_nlm(int64_t P, int L, const double* rnlm, double* nlm, hipStream_t s) {
    if (P == 0) return hipSuccess;
    int n = (L + 1) * (L + 2) / 2, rn = rnlm_len(L);
    hipLaunchKernelGGL(rnlm_to_nlm_kernel, dim3(nb(P * n, 256)), dim3(256), 0, s, P, n, rn, rnlm, nlm);
    return hipGetLastError();
}

}  // namespace sfg
