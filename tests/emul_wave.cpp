// CPU emulation driver (tests only) for the wave-velocity and full-a4
// kernels: builds the same host tables as the library (rotation blocks,
// a4 maps, elastic coefficients) and runs the kernels lane by lane.
#include <cstdint>
#include <vector>

#include "sfg_internal.h"
#include "sfg_wave.h"

thread_local sfg_dim3 threadIdx, blockIdx, blockDim, gridDim;
void sfg_emul_syncthreads() {}

namespace sfg {
__global__ void vi_elastic_kernel(ViElasticArgs a);
__global__ void vi_em_kernel(ViEMArgs a);
__global__ void a4full_kernel(A4FullArgs a);
template <bool CART, int MODE>
__global__ void eij_ortho_kernel(OrthoArgs a);
template <int MODE>
__global__ void vi_ortho_kernel(OrthoArgs a);
}

template <class F>
static void grid(int64_t N, int bs, F f) {
    for (int64_t b = 0; b < (N + bs - 1) / bs; b++)
        for (int t = 0; t < bs; t++) {
            blockIdx.x = (unsigned)b;
            blockDim.x = (unsigned)bs;
            gridDim.x = (unsigned)((N + bs - 1) / bs);
            threadIdx.x = (unsigned)t;
            f();
        }
}

namespace {
struct Tables {
    std::vector<double> A2, A4, A4F, R;
    int nq;
};
Tables tables(int L, int ndir, const double* theta, const double* phi) {
    const sfg::RealOperator& op = sfg::real_operator(L);
    Tables t;
    t.nq = op.n < 15 ? op.n : 15;
    t.A2.resize(6 * 15);
    t.A4.resize(36 * 15);
    t.A4F.resize(15 * 15);
    for (int i = 0; i < 6; i++) for (int q = 0; q < 15; q++) t.A2[i * 15 + q] = op.A2[i][q];
    for (int i = 0; i < 36; i++) for (int q = 0; q < 15; q++) t.A4[i * 15 + q] = op.A4[i][q];
    double m[15][15];
    sfg::a4full_map(m);
    for (int e = 0; e < 15; e++) for (int q = 0; q < 15; q++) t.A4F[e * 15 + q] = q < t.nq ? m[e][q] : 0.0;
    t.R.resize((size_t)ndir * sfg::ROT4_STRIDE);
    for (int d = 0; d < ndir; d++) {
        double R2[5][5], R4[9][9];
        sfg::rotation_l4(-theta[d], -phi[d], R2, R4);
        double* o = t.R.data() + (size_t)d * sfg::ROT4_STRIDE;
        for (int i = 0; i < 5; i++) for (int j = 0; j < 5; j++) o[5 * i + j] = R2[i][j];
        for (int i = 0; i < 9; i++) for (int j = 0; j < 9; j++) o[25 + 9 * i + j] = R4[i][j];
    }
    return t;
}
}  // namespace

// x: real SoA state (n, P).  Vi: (P, ndir, 3).
extern "C" int emul_vi_elastic(int L, int64_t P, const double* x, double alpha, const double* Lame, double rho,
                               int ndir, const double* theta, const double* phi, double* Vi, int* status) {
    Tables t = tables(L, ndir, theta, phi);
    sfg::ViElasticArgs a{};
    a.P = P; a.ndir = ndir; a.nq = t.nq; a.x = x; a.A2 = t.A2.data(); a.A4 = t.A4.data(); a.A4F = t.A4F.data();
    a.R = t.R.data(); a.alpha = alpha; a.rho = rho; a.Vi = Vi; a.status = status;
    sfg::elas_params(Lame, a.kf, a.kr);
    grid(P * ndir, 128, [&] { sfg::vi_elastic_kernel(a); });
    return 0;
}

// the same with the compliance and stiffness coefficients given directly
// (kf, kr of elas_params), to reach branches no grain parameters reach
extern "C" int emul_vi_elastic_k(int L, int64_t P, const double* x, double alpha, const double* kf,
                                 const double* kr, double rho, int ndir, const double* theta, const double* phi,
                                 double* Vi, int* status) {
    Tables t = tables(L, ndir, theta, phi);
    sfg::ViElasticArgs a{};
    a.P = P; a.ndir = ndir; a.nq = t.nq; a.x = x; a.A2 = t.A2.data(); a.A4 = t.A4.data(); a.A4F = t.A4F.data();
    a.R = t.R.data(); a.alpha = alpha; a.rho = rho; a.Vi = Vi; a.status = status;
    for (int i = 0; i < 5; i++) { a.kf[i] = kf[i]; a.kr[i] = kr[i]; }
    grid(P * ndir, 128, [&] { sfg::vi_elastic_kernel(a); });
    return 0;
}

extern "C" int emul_vi_em(int L, int64_t P, const double* x, double eps_m, double eps_t, double mu, int ndir,
                          const double* theta, const double* phi, double* Vi) {
    Tables t = tables(L, ndir, theta, phi);
    sfg::ViEMArgs a{};
    a.P = P; a.ndir = ndir; a.x = x; a.A2 = t.A2.data(); a.R = t.R.data();
    a.eps_m = eps_m; a.eps_t = eps_t; a.mu = mu; a.Vi = Vi;
    grid(P * ndir, 128, [&] { sfg::vi_em_kernel(a); });
    return 0;
}

extern "C" int emul_a4(int L, int64_t P, const double* x, double* a4) {
    double th = 0, ph = 0;
    Tables t = tables(L, 1, &th, &ph);
    sfg::A4FullArgs a{};
    a.P = P; a.nq = t.nq; a.x = x; a.A4F = t.A4F.data(); a.a4 = a4;
    grid(P, 128, [&] { sfg::a4full_kernel(a); });
    return 0;
}

// orthotropic grains: x1, x2, x3 real SoA states (n, P) (x3 null: v = b x n);
// frames (P, 3, 3) or null.  Eij: (P, 6); Vi: (P, ndir, 3).
extern "C" int emul_eij_ortho(int L, int64_t P, const double* x1, const double* x2, const double* x3,
                              const double* frames, const double* ci, double* Eij) {
    const sfg::RealOperator& op = sfg::real_operator(L);
    sfg::OrthoArgs a{};
    a.P = P; a.nq = op.n < 15 ? op.n : 15; a.x1 = x1; a.x2 = x2; a.x3 = x3; a.frames = frames; a.Eij = Eij;
    for (int i = 0; i < 6; i++) a.ci[i] = ci[i];
    sfg::ortho_q6_iso(a.ci, a.q6iso, &a.dqiso);
    if (frames) {
        if (x3) grid(P, 256, [&] { sfg::eij_ortho_kernel<false, 0>(a); });
        else grid(P, 256, [&] { sfg::eij_ortho_kernel<false, 1>(a); });
    } else {
        if (x3) grid(P, 256, [&] { sfg::eij_ortho_kernel<true, 0>(a); });
        else grid(P, 256, [&] { sfg::eij_ortho_kernel<true, 1>(a); });
    }
    return 0;
}

extern "C" int emul_vi_ortho(int L, int64_t P, const double* x1, const double* x2, const double* x3,
                             const double* Lame, double rho, int ndir, const double* theta, const double* phi,
                             double* Vi) {
    Tables t = tables(L, ndir, theta, phi);
    sfg::OrthoArgs a{};
    a.P = P; a.nq = t.nq; a.ndir = ndir; a.x1 = x1; a.x2 = x2; a.x3 = x3; a.R = t.R.data(); a.rho = rho; a.Vi = Vi;
    for (int i = 0; i < 6; i++) a.lam[i] = Lame[i];
    for (int i = 0; i < 3; i++) a.mu[i] = Lame[6 + i];
    // one workgroup per (parcel block, direction): the kernel derives both from blockIdx.x
    const int64_t nw = (P + 255) / 256 * ndir;
    if (x3) grid(nw * 256, 256, [&] { sfg::vi_ortho_kernel<0>(a); });
    else grid(nw * 256, 256, [&] { sfg::vi_ortho_kernel<1>(a); });
    return 0;
}
