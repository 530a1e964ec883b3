// CPU emulation driver (tests only) for the rotation and a4-eigentensor
// kernels: builds the library's host tables (Wigner sum-formula terms, a2 and
// full-a4 maps) and runs the kernels lane by lane.
#include <cstdint>
#include <vector>

#include "sfg_internal.h"
#include "sfg_rotate.h"

thread_local sfg_dim3 threadIdx, blockIdx, blockDim, gridDim;
void sfg_emul_syncthreads() {}

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

// x, out: real SoA states (n, P); theta, phi (P)
extern "C" int emul_rotate(int L, int64_t P, const double* x, const double* theta, const double* phi, double* out) {
    std::vector<sfg::WignerTerm> terms;
    std::vector<int> first;
    sfg::wigner_terms(sfg::ROT_LMAX, terms, first);
    if ((int)terms.size() != sfg::wigner_offset(sfg::ROT_LMAX + 2, 0, -(sfg::ROT_LMAX + 2))) return -1;
    std::vector<double> coef;
    for (auto& t : terms) coef.push_back(t.coef);
    sfg::RotateArgs a{};
    a.P = P; a.L = L; a.x = x; a.out = out; a.theta = theta; a.phi = phi; a.coef = coef.data();
    switch (L) {
        case 2: grid(P, 128, [&] { sfg::rotate_kernel<2>(a); }); break;
        case 4: grid(P, 128, [&] { sfg::rotate_kernel<4>(a); }); break;
        case 6: grid(P, 128, [&] { sfg::rotate_kernel<6>(a); }); break;
        case 8: grid(P, 128, [&] { sfg::rotate_kernel<8>(a); }); break;
        case 10: grid(P, 128, [&] { sfg::rotate_kernel<10>(a); }); break;
        case 12: grid(P, 128, [&] { sfg::rotate_kernel<12>(a); }); break;
        default: return -1;
    }
    return 0;
}

// x: real SoA states (n, P), n >= 15; Q (P, 6, 3, 3), ev (P, 6), valid (P)
extern "C" int emul_a4eig(int L, int64_t P, const double* x, double* Q, double* ev, int* valid) {
    const sfg::RealOperator& op = sfg::real_operator(L);
    std::vector<double> A2(6 * 15), A4F(15 * 15);
    for (int i = 0; i < 6; i++) for (int q = 0; q < 15; q++) A2[i * 15 + q] = op.A2[i][q];
    double m[15][15];
    sfg::a4full_map(m);
    for (int e = 0; e < 15; e++) for (int q = 0; q < 15; q++) A4F[e * 15 + q] = m[e][q];
    sfg::A4EigArgs a{};
    a.P = P; a.x = x; a.A2 = A2.data(); a.A4F = A4F.data(); a.Q = Q; a.ev = ev; a.valid = valid;
    grid(P, 128, [&] { sfg::a4eig_kernel(a); });
    return 0;
}
