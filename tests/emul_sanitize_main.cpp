// Sanitizer run of the CPU emulation (tests only): built together with
// emul_driver.cpp, the parameter kernel, the generic stepper and one generated
// stepper under AddressSanitizer and UndefinedBehaviorSanitizer (host code;
// see test_emul_kernels.py), it steps a seeded batch of one full and one
// partial wavefront with both steppers and checks that the two agree and that
// no parcel is flagged.  Exit status 0 on success; a sanitizer report aborts.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

extern "C" int emul_evolve(int L, int64_t P, int procs, int scheme, int nsteps, double dt, double iota, double zeta,
                           const double* ugrad, const double* gam0, const double* lam, double* x, int which,
                           int* status);

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s L scheme(0 euler, 1 rk4)\n", argv[0]);
        return 2;
    }
    const int L = std::atoi(argv[1]), scheme = std::atoi(argv[2]);
    const int64_t P = 70;
    int n = 0;
    for (int l = 0; l <= L; l += 2) n += 2 * l + 1;   // real coefficients: m = 0 and (re, im) of m > 0
    std::mt19937_64 rng(1234 + L);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::uniform_real_distribution<double> ud(0.0, 1.0);
    std::vector<double> x0((size_t)n * P), ug((size_t)9 * P), g0(P), lam(P);
    // perturbed isotropic states: x_0 = 1/sqrt(4 pi), small higher coefficients
    for (int64_t p = 0; p < P; p++) {
        x0[p] = 0.28209479177387814;
        for (int a = 1; a < n; a++) x0[(size_t)a * P + p] = 1e-2 * nd(rng) * 0.28209479177387814;
        double t = 0;
        for (int i = 0; i < 9; i++) ug[(size_t)p * 9 + i] = 1e-3 * nd(rng);
        for (int i = 0; i < 3; i++) t += ug[(size_t)p * 9 + 4 * i];
        for (int i = 0; i < 3; i++) ug[(size_t)p * 9 + 4 * i] -= t / 3;   // traceless
        g0[p] = std::pow(10.0, -4 + 2 * ud(rng));
        lam[p] = std::pow(10.0, -4 + ud(rng));
    }
    const int procs = 1 | 2 | 4 | 8;   // LROT, DDRX, CDRX, REG
    std::vector<double> xa = x0, xb = x0;
    std::vector<int> sa(P, 0), sb(P, 0);
    emul_evolve(L, P, procs, scheme, 3, 0.5, 1.0, 0.0, ug.data(), g0.data(), lam.data(), xa.data(), 0, sa.data());
    emul_evolve(L, P, procs, scheme, 3, 0.5, 1.0, 0.0, ug.data(), g0.data(), lam.data(), xb.data(), 1, sb.data());
    double err = 0, mx = 0;
    for (size_t i = 0; i < xa.size(); i++) {
        err = std::fmax(err, std::fabs(xa[i] - xb[i]));
        mx = std::fmax(mx, std::fabs(xb[i]));
    }
    int flagged = 0;
    for (int64_t p = 0; p < P; p++) flagged += (sa[p] != 0) + (sb[p] != 0);
    std::printf("L=%d scheme=%d generated vs generic max abs diff %.3e (scale %.3e), flagged %d\n", L, scheme, err, mx,
                flagged);
    return (err <= 1e-12 * mx && flagged == 0) ? 0 : 1;
}
