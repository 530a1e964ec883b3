// HIP kernels of libspecfab_gpu (gfx950).
//
//  * params_kernel      per-parcel operator parameters (LROT/DDRX weights,
//                       <D> weights, REG/CDRX diagonal) from the kinematics
//  * evolve_generic     table-driven stepping for any L <= 20 (one lane per
//                       parcel, state in HBM, coupling terms via scalar loads)
//  * conversions, moments and the dense per-parcel operator builders.
//
// The register-resident fast steppers (generated per L) live in
// sfg_step_fast.hip.
#include "sfg_kernels.h"
#ifndef SFG_HOST_EMUL
#include "sfg_fast.h"
#endif
#include "sfg_device.h"

namespace sfg {

// ------------------------------------------------------------------ params

__global__ void params_kernel(ParamArgs a) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.P) return;
    const int64_t P = a.P;
    double G[9], D[9], W[9], tau[9];
    for (int i = 0; i < 9; i++) G[i] = a.ugrad[p * 9 + i];
    sym_skew(G, D, W);
    for (int i = 0; i < 9; i++) tau[i] = a.tau ? a.tau[p * 9 + i] : D[i];
    double th[NPAR_K];
    for (int i = 0; i < NPAR_K; i++) th[i] = 0.0;
    if (a.procs & 1) {
        cplx qe[5], qo[3];
        d_lrot_weights(D, W, a.iota, a.zeta, nullptr, nullptr, nullptr, nullptr, qe, qo);
        th[0] = qe[2].re; th[1] = qe[3].re; th[2] = qe[3].im; th[3] = qe[4].re; th[4] = qe[4].im;
        th[5] = qo[1].re; th[6] = qo[2].re; th[7] = qo[2].im;
    }
    double om[15];
    for (int i = 0; i < 15; i++) om[i] = 0.0;
    if (a.procs & 2) {
        double g0 = a.gam0 ? a.gam0[p] : 0.0;
        cplx g[15];
        d_ddrx_weights(tau, g);
        // theta = Gamma0 * (Re/Im g(l, m >= 0)), k index of (l,m): l=0 -> 0, l=2 -> 3+m, l=4 -> 10+m
        th[8] = g0 * g[0].re;
        th[9] = g0 * g[3].re;
        th[10] = g0 * g[4].re; th[11] = g0 * g[4].im;
        th[12] = g0 * g[5].re; th[13] = g0 * g[5].im;
        th[14] = g0 * g[10].re;
        for (int m = 1; m <= 4; m++) {
            th[13 + 2 * m] = g0 * g[10 + m].re;
            th[14 + 2 * m] = g0 * g[10 + m].im;
        }
        // <D> = 5 [ (tau.tau):a2 - tau:a4:tau ] / (tau:tau), src/dynamics.f90:402-422,
        // with a2v = A2 x / x0 + I/3, a4v = A4 x / x0  =>  Gamma0 <D> = (om . x) / x0
        const double s2 = sqrt(2.0);
        double tsq[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double s = 0;
                for (int k = 0; k < 3; k++) s += tau[3 * i + k] * tau[3 * k + j];
                tsq[3 * i + j] = s;
            }
        double norm = tsq[0] + tsq[4] + tsq[8];
        double tv[6] = {tau[0], tau[4], tau[8], s2 * tau[5], s2 * tau[2], s2 * tau[1]};
        double sv[6] = {tsq[0], tsq[4], tsq[8], s2 * tsq[5], s2 * tsq[2], s2 * tsq[1]};
        double c = g0 * 5.0 / norm;
        for (int q = 0; q < a.nq; q++) {
            double d1 = 0, d2 = 0;
            for (int i = 0; i < 6; i++) {
                d1 += sv[i] * a.A2[i * 15 + q];
                double r = 0;
                for (int j = 0; j < 6; j++) r += a.A4[(6 * i + j) * 15 + q] * tv[j];
                d2 += tv[i] * r;
            }
            om[q] = c * (d1 - d2);
        }
        om[0] += c * ((sv[0] + sv[1] + sv[2]) / 3.0);
    }
    for (int i = 0; i < NPAR_K; i++) a.theta[(int64_t)i * P + p] = th[i];
    for (int i = 0; i < 15; i++) a.omega[(int64_t)i * P + p] = om[i];
    double ratemag = 0.0;
    if (a.procs & 8) {
        double fro = 0;
        for (int i = 0; i < 9; i++) fro += D[i] * D[i];
        ratemag = a.reg_nu * sqrt(fro);
    }
    double lam = ((a.procs & 4) && a.lam) ? a.lam[p] : 0.0;
    for (int b = 0; b < a.nblk; b++) {
        int l = 2 * b;
        double v = -ratemag * a.reg_shape[b];
        if (a.procs & 4) v += lam * (-(double)(l * (l + 1)));
        a.sl[(int64_t)b * P + p] = v;
    }
}

// ------------------------------------------------------------------ generic stepper

// f = R(theta) x + (s_l - Gamma0 <D>(x)) x for all rows, written through a
// per-row callback.  `in` is an SoA array (row stride P).
struct RowCtx {
    const EvolveArgs* a;
    int64_t p;
};

template <class Emit>
__device__ __forceinline__ void rate_generic(const EvolveArgs& a, int64_t p, const double* in, Emit emit) {
    const int64_t P = a.P;
    double gD = 0.0;
    if (a.use_ddrx) {
        double s = 0;
        for (int q = 0; q < a.nq; q++) s += a.omega[(int64_t)q * P + p] * in[(int64_t)q * P + p];
        gD = s / in[p];
    }
    for (int r = 0; r < a.n; r++) {
        int2 rg = a.rows[r];
        double acc = 0.0;
        for (int g = rg.x; g < rg.y; g++) {
            int2 tr = a.grp_rng[g];
            double u = 0.0;
            for (int t = tr.x; t < tr.y; t++) u = fma(a.tval[t], in[(int64_t)a.tcol[t] * P + p], u);
            acc = fma(a.theta[(int64_t)a.grp_param[g] * P + p], u, acc);
        }
        double xr = in[(int64_t)r * P + p];
        double f = fma(a.sl[(int64_t)a.row_blk[r] * P + p] - gD, xr, acc);
        emit(r, f);
    }
}

__global__ void evolve_generic(EvolveArgs a) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.P) return;
    const int64_t P = a.P;
    const int rem = a.steps ? a.steps[p] : a.nsteps;
    const int ns = (a.steps && a.steps_dec) ? (rem < a.nsteps ? rem : a.nsteps) : rem;
    const double h = a.dt;
    double* X = a.X;
    double* YS = a.YS;
    double* XA = a.XA;
    double* XB = a.XB;
    for (int step = 0; step < ns; step++) {
        if (a.scheme == 0) {
            rate_generic(a, p, X, [&](int r, double f) { XA[(int64_t)r * P + p] = fma(h, f, X[(int64_t)r * P + p]); });
            for (int r = 0; r < a.n; r++) X[(int64_t)r * P + p] = XA[(int64_t)r * P + p];
        } else {
            // stage 1: in = X
            rate_generic(a, p, X, [&](int r, double f) {
                double x0 = X[(int64_t)r * P + p];
                YS[(int64_t)r * P + p] = fma(h / 6, f, x0);
                XA[(int64_t)r * P + p] = fma(h / 2, f, x0);
            });
            // stage 2: in = XA
            rate_generic(a, p, XA, [&](int r, double f) {
                double x0 = X[(int64_t)r * P + p];
                YS[(int64_t)r * P + p] = fma(h / 3, f, YS[(int64_t)r * P + p]);
                XB[(int64_t)r * P + p] = fma(h / 2, f, x0);
            });
            // stage 3: in = XB
            rate_generic(a, p, XB, [&](int r, double f) {
                double x0 = X[(int64_t)r * P + p];
                YS[(int64_t)r * P + p] = fma(h / 3, f, YS[(int64_t)r * P + p]);
                XA[(int64_t)r * P + p] = fma(h, f, x0);
            });
            // stage 4: in = XA, out = X
            rate_generic(a, p, XA, [&](int r, double f) {
                X[(int64_t)r * P + p] = fma(h / 6, f, YS[(int64_t)r * P + p]);
            });
        }
    }
    if (a.steps && a.steps_dec) a.steps[p] = rem - ns;
    if (a.status) {
        int st = 0;
        for (int r = 0; r < a.n; r++) {
            double v = X[(int64_t)r * P + p];
            if (!isfinite(v)) st |= 1;
        }
        a.status[p] |= st;
    }
}

// ------------------------------------------------------------------ conversions

// nlm (P, n) complex interleaved -> x (n, P) real SoA
__global__ void nlm_to_state_kernel(int64_t P, int n, const int* real_to_cplx, const int* real_part,
                                    const double* nlm, double* x) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= P * n) return;
    int64_t p = idx % P;
    int a = (int)(idx / P);
    int j = real_to_cplx[a];
    x[(int64_t)a * P + p] = nlm[(p * n + j) * 2 + real_part[a]];
}

// x (n, P) -> nlm (P, n) complex interleaved, n_l^-m = (-1)^m conj(n_l^m)
__global__ void state_to_nlm_kernel(int64_t P, int n, const int* cre, const int* cim, const int* csgn,
                                    const int* cneg, const double* x, double* nlm) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= P * n) return;
    int j = (int)(idx % n);
    int64_t p = idx / n;
    double re = x[(int64_t)cre[j] * P + p];
    double im = cim[j] >= 0 ? x[(int64_t)cim[j] * P + p] : 0.0;
    if (cneg[j]) im = -im;
    double s = csgn[j];
    nlm[idx * 2] = s * re;
    nlm[idx * 2 + 1] = s * im;
}

// ------------------------------------------------------------------ moments

// a2 (P,3,3) row-major from the real state: src/include/ev_c2__body.f90
__global__ void a2_kernel(int64_t P, const double* x, double* a2) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    double n00 = x[p];
    // real indices of l=2: 1 = Re n20, 2 = Re n21, 3 = Im n21, 4 = Re n22, 5 = Im n22
    double h0 = x[1 * P + p] / n00, h1r = x[2 * P + p] / n00, h1i = x[3 * P + p] / n00;
    double h2r = x[4 * P + p] / n00, h2i = x[5 * P + p] / n00;
    const double c = sqrt(2.0 / 3), s = sqrt(2 / 15.0);
    double e[9];
    e[0] = -0.5 * c * h0 + h2r;
    e[4] = -0.5 * c * h0 - h2r;
    e[8] = c * h0;
    e[1] = e[3] = -h2i;
    e[2] = e[6] = -h1r;
    e[5] = e[7] = h1i;
    for (int i = 0; i < 9; i++) a2[p * 9 + i] = s * e[i] + ((i % 4 == 0) ? 1.0 / 3.0 : 0.0);
}

// a4 Mandel (P,6,6) via the linear map A4 (a4v = A4 x / x0)
__global__ void a4mandel_kernel(int64_t P, int nq, const double* A4, const double* x, double* a4v) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    double xq[15];
    for (int q = 0; q < 15; q++) xq[q] = q < nq ? x[(int64_t)q * P + p] : 0.0;
    for (int i = 0; i < 36; i++) {
        double s = 0;
        for (int q = 0; q < nq; q++) s += A4[i * 15 + q] * xq[q];
        a4v[p * 36 + i] = s / xq[0];
    }
}

// <D> (src/dynamics.f90:402-422) of real states x (n, P) for stresses tau (P,3,3)
__global__ void evD2_kernel(int64_t P, int nq, const double* A2, const double* A4, const double* x, const double* tau,
                            double* out) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    double xq[15];
    for (int q = 0; q < 15; q++) xq[q] = q < nq ? x[(int64_t)q * P + p] : 0.0;
    double a2v[6], a4v[36];
    for (int i = 0; i < 6; i++) {
        double s = 0;
        for (int q = 0; q < nq; q++) s += A2[i * 15 + q] * xq[q];
        a2v[i] = s / xq[0] + (i < 3 ? 1.0 / 3.0 : 0.0);
    }
    for (int i = 0; i < 36; i++) {
        double s = 0;
        for (int q = 0; q < nq; q++) s += A4[i * 15 + q] * xq[q];
        a4v[i] = s / xq[0];
    }
    const double* t = tau + p * 9;
    const double s2 = sqrt(2.0);
    double tsq[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += t[3 * i + k] * t[3 * k + j];
            tsq[3 * i + j] = s;
        }
    double norm = tsq[0] + tsq[4] + tsq[8];
    double tv[6] = {t[0], t[4], t[8], s2 * t[5], s2 * t[2], s2 * t[1]};
    double sv[6] = {tsq[0], tsq[4], tsq[8], s2 * tsq[5], s2 * tsq[2], s2 * tsq[1]};
    double d1 = 0, d2 = 0;
    for (int i = 0; i < 6; i++) {
        d1 += sv[i] * a2v[i];
        double r = 0;
        for (int j = 0; j < 6; j++) r += a4v[6 * i + j] * tv[j];
        d2 += tv[i] * r;
    }
    out[p] = 5 * (d1 - d2) / norm;
}

// Diagonal operators as dense real (P, n, n): mode 0 = M_CDRX, 1 = M_REG(D)
__global__ void diag_op_kernel(int64_t P, int n, int mode, const int* row_l, const double* reg_shape, double reg_nu,
                               const double* D, double* M) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= P * n * n) return;
    int64_t p = idx / ((int64_t)n * n);
    int r = (int)(idx % ((int64_t)n * n));
    int i = r / n, j = r % n;
    double v = 0.0;
    if (i == j) {
        int l = row_l[i];
        if (mode == 0) {
            v = -(double)(l * (l + 1));
        } else {
            double fro = 0;
            for (int q = 0; q < 9; q++) fro += D[p * 9 + q] * D[p * 9 + q];
            v = -(reg_nu * sqrt(fro)) * reg_shape[l / 2];
        }
    }
    M[idx] = v;
}

// ------------------------------------------------------------------ dense operators

// Complex per-parcel operator, row-major (P, n, n), from the coupling-table
// CSR (entries of row i sorted by j): one thread per (parcel, row).
//   mode 0: M_LROT   = -sum GC g0 + GCm gz + GC_m1 gn + GC_p1 gp  (k < 6)
//   mode 1: M_DDRX_src = sum GC g                                  (k < 15)
// One workgroup of four waves per node: thread 0 evaluates the node's
// coupling weights once into LDS, then wave q builds rows q, q + 4, ..., lane
// c owning columns c, c + 64, ...: it sums the entries of (row, column) in
// table order and stores the element once, so every row leaves as contiguous
// 1 KiB runs and nothing is read back.
// the node's coupling weights into sw (one thread)
__device__ inline void dense_weights(const DenseArgs& a, int64_t p, cplx (*sw)[15]) {
    {
        cplx w[4][15];
        if (a.mode == 0) {
            double D[9], W[9];
            for (int q = 0; q < 9; q++) { D[q] = a.A[p * 9 + q]; W[q] = a.B[p * 9 + q]; }
            cplx qe[5], qo[3];
            d_lrot_weights(D, W, a.iota, a.zeta, w[0], w[1], w[2], w[3], qe, qo);
            for (int t = 0; t < 4; t++)
                for (int k = 0; k < 6; k++) sw[t][k] = w[t][k];
        } else {
            double T[9];
            for (int q = 0; q < 9; q++) T[q] = a.A[p * 9 + q];
            d_ddrx_weights(T, w[0]);
            for (int k = 0; k < 15; k++) sw[0][k] = w[0][k];
        }
    }
}

// element (i, c) of the node's operator: its table entries in table order
__device__ inline void dense_element(const DenseArgs& a, const cplx (*sw)[15], int i, int c, double& re, double& im) {
    const int kmax = a.mode == 0 ? 6 : 15;
    const double sg = a.mode == 0 ? -1.0 : 1.0;
    const int2 r = a.seg[(size_t)i * a.segw + c];
    re = 0.0;
    im = 0.0;
    for (int e = r.x; e < r.y; e++) {
        int4 ent = a.ent[e];   // (t, j = c, k, -)
        int t = ent.x, k = ent.z;
        if (k >= kmax) continue;
        if (a.mode == 1 && t != 0) continue;
        double v = a.val[e];
        re += sg * v * sw[t][k].re;
        im += sg * v * sw[t][k].im;
    }
}

__global__ __launch_bounds__(256) void dense_op_kernel(DenseArgs a) {
    __shared__ cplx sw[4][15];
    const int64_t p = blockIdx.x;
    if (threadIdx.x == 0) dense_weights(a, p, sw);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    for (int i = threadIdx.x >> 6; i < a.n; i += 4) {
        double* row = a.M + (p * a.n + i) * a.n * 2;
        for (int c = lane; c < a.n; c += 64) {
            double re, im;
            dense_element(a, sw, i, c, re, im);
            row[2 * c] = re;
            row[2 * c + 1] = im;
        }
    }
}

// the coupling weights of every node into W (P, 48): real parts of the K
// components, then (at +24) their imaginary parts
__global__ __launch_bounds__(256) void dense_weights_kernel(DenseArgs a, double* W) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.P) return;
    cplx sw[4][15];
    dense_weights(a, p, sw);
    double* o = W + p * RMAP_W;
    if (a.mode == 0) {
        for (int t = 0; t < 4; t++)
            for (int k = 0; k < 6; k++) {
                o[t * 6 + k] = sw[t][k].re;
                o[24 + t * 6 + k] = sw[t][k].im;
            }
    } else {
        for (int k = 0; k < 15; k++) {
            o[k] = sw[0][k].re;
            o[24 + k] = sw[0][k].im;
        }
    }
}

// block entries r of the nodes of RMAP_NODES-node chunks, one entry per
// thread: entry r's S slots (weight component, two coefficients) in
// registers, the chunk's weights staged in LDS, per node nt (wave-uniform)
// terms and four coalesced stores.  The terms skipped are exact zeros and
// the rest keep their order, so each sum equals the dense sum over all K
// components.
template <int S>
__global__ __launch_bounds__(256) void reduced_map_kernel(int64_t P, RMap m, const double* __restrict__ W,
                                                          double* Mrr, double* Mri, double* Mir, double* Mii) {
    __shared__ double ws[RMAP_NODES * RMAP_W];
    const int rn2 = m.rn2;
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = r < rn2;
    // waves past the last entry have no trips; a partial last wave reads its
    // own (in-range) count
    const int nt = (r >> 6) <= (rn2 - 1) >> 6 ? m.nt[r >> 6] : 0;
    int ci[S];
    double cp[S], cm[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const bool use = on && s < nt;
        ci[s] = use ? m.ci[s * rn2 + r] : 0;
        cp[s] = use ? m.cp[s * rn2 + r] : 0.0;
        cm[s] = use ? m.cm[s * rn2 + r] : 0.0;
    }
    const int64_t nchunk = (P + RMAP_NODES - 1) / RMAP_NODES;
    for (int64_t ch = blockIdx.y; ch < nchunk; ch += gridDim.y) {
        const int64_t p0 = ch * RMAP_NODES;
        const int np = (int)(P - p0 < RMAP_NODES ? P - p0 : RMAP_NODES);
        __syncthreads();   // the previous chunk's weights are no longer read
        for (int i = threadIdx.x; i < np * RMAP_W; i += blockDim.x) ws[i] = W[p0 * RMAP_W + i];
        __syncthreads();
        for (int q = 0; q < np; q++) {
            const double* w = ws + q * RMAP_W;
            double rr = 0.0, ir = 0.0, ii = 0.0, ri = 0.0;
#pragma unroll
            for (int s = 0; s < S; s++) {
                if (s < nt) {
                    const double wr = w[ci[s]], wi = w[24 + ci[s]];
                    rr += cp[s] * wr;
                    ir += cp[s] * wi;
                    ii += cm[s] * wr;
                    ri += cm[s] * wi;
                }
            }
            if (on) {
                const int64_t t = (p0 + q) * rn2 + r;
                Mrr[t] = rr;
                Mri[t] = -ri;
                Mir[t] = ir;
                Mii[t] = ii;
            }
        }
    }
}

// ------------------------------------------------------------------ launchers
#ifndef SFG_HOST_EMUL

hipError_t launch_dense_weights(const DenseArgs& a, double* W, hipStream_t s) {
    if (a.P == 0) return hipSuccess;
    hipLaunchKernelGGL(dense_weights_kernel, dim3((unsigned)((a.P + 255) / 256)), dim3(256), 0, s, a, W);
    return hipGetLastError();
}

hipError_t launch_reduced_map(int64_t P, const RMap& m, const double* W, double* Mrr, double* Mri, double* Mir,
                              double* Mii, hipStream_t s) {
    if (P == 0) return hipSuccess;
    const int64_t nchunk = (P + RMAP_NODES - 1) / RMAP_NODES;
    const dim3 g((unsigned)((m.rn2 + 255) / 256), (unsigned)(nchunk < 65535 ? nchunk : 65535)), b(256);
    if (m.S == 10)
        hipLaunchKernelGGL(reduced_map_kernel<10>, g, b, 0, s, P, m, W, Mrr, Mri, Mir, Mii);
    else if (m.S == 5)
        hipLaunchKernelGGL(reduced_map_kernel<5>, g, b, 0, s, P, m, W, Mrr, Mri, Mir, Mii);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

static inline unsigned nblocks(int64_t work, int bs) { return (unsigned)((work + bs - 1) / bs); }

hipError_t launch_params(const ParamArgs& a, hipStream_t s) {
    if (a.P == 0) return hipSuccess;
    hipLaunchKernelGGL(params_kernel, dim3(nblocks(a.P, 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

const void* generic_kernel() { return (const void*)&evolve_generic; }

hipError_t launch_evolve_generic(const EvolveArgs& a, hipStream_t s) {
    if (a.P == 0) return hipSuccess;
    hipLaunchKernelGGL(evolve_generic, dim3(nblocks(a.P, 64)), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_nlm_to_state(int64_t P, int n, const int* r2c, const int* rpart, const double* nlm, double* x,
                               hipStream_t s) {
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(nlm_to_state_kernel, dim3(nblocks(P * n, 256)), dim3(256), 0, s, P, n, r2c, rpart, nlm, x);
    return hipGetLastError();
}

hipError_t launch_state_to_nlm(int64_t P, int n, const int* cre, const int* cim, const int* csgn, const int* cneg,
                               const double* x, double* nlm, hipStream_t s) {
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(state_to_nlm_kernel, dim3(nblocks(P * n, 256)), dim3(256), 0, s, P, n, cre, cim, csgn, cneg,
                       x, nlm);
    return hipGetLastError();
}

hipError_t launch_a2(int64_t P, const double* x, double* a2, hipStream_t s) {
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(a2_kernel, dim3(nblocks(P, 256)), dim3(256), 0, s, P, x, a2);
    return hipGetLastError();
}

hipError_t launch_a4mandel(int64_t P, int nq, const double* A4, const double* x, double* a4v, hipStream_t s) {
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(a4mandel_kernel, dim3(nblocks(P, 256)), dim3(256), 0, s, P, nq, A4, x, a4v);
    return hipGetLastError();
}

hipError_t launch_evD2(int64_t P, int nq, const double* A2, const double* A4, const double* x, const double* tau,
                       double* out, hipStream_t s) {
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(evD2_kernel, dim3(nblocks(P, 256)), dim3(256), 0, s, P, nq, A2, A4, x, tau, out);
    return hipGetLastError();
}

hipError_t launch_diag_op(int64_t P, int n, int mode, const int* row_l, const double* reg_shape, double reg_nu,
                          const double* D, double* M, hipStream_t s) {
    if (P == 0) return hipSuccess;
    hipLaunchKernelGGL(diag_op_kernel, dim3(nblocks(P * n * n, 256)), dim3(256), 0, s, P, n, mode, row_l, reg_shape,
                       reg_nu, D, M);
    return hipGetLastError();
}

hipError_t launch_dense_op(const DenseArgs& a, hipStream_t s) {
    if (a.P == 0) return hipSuccess;
    if (a.P > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(dense_op_kernel, dim3((unsigned)a.P), dim3(256), 0, s, a);
    return hipGetLastError();
}

#endif  // SFG_HOST_EMUL

}  // namespace sfg
