// Host-side tables of the wave-velocity path: the full fourth-order structure
// tensor as a linear map of the real l <= 4 state, and the Wigner-D rotation
// of that block in the real basis, both with the reference's REAL(4) literals.
#include <cmath>
#include <complex>

#include "sfg_internal.h"

namespace sfg {

namespace {

using cd = std::complex<double>;

// One term c * n_l^m of an a4 entry; c is formed in float32 exactly as the
// reference literal: kind 0 -> a, 1 -> a*sqrt(b), 2 -> a*b**1.5.
struct A4Term { int l, m, kind; float a, b; };
// Outer factor of an entry: 1, i (complex(4) (0,1)), or i*Sqrt((2.0)).
enum { OUT_1 = 0, OUT_I = 1, OUT_IS2 = 2 };
struct A4Entry { int outer, div5, nterm; A4Term t[9]; };

// The 15 index-sorted entries (1111, 1112, 1113, 1122, 1123, 1133, 1222, 1223,
// 1233, 1333, 2222, 2223, 2233, 2333, 3333) of src/include/ev_c4__body.f90,
// a Mathematica export whose literals are REAL(4).  The coefficients have to
// be the reference's float32 products (e.g. (-12.0)*Sqrt((5.0)), (3.0)**1.5)
// for FP64 parity: exact constants differ from them by ~3e-8 relative.
const A4Entry kA4[15] = {
    {OUT_1, 1, 9, {{0, 0, 0, 42, 0}, {2, 0, 1, -12, 5}, {2, -2, 1, 6, 30}, {2, 2, 1, 6, 30}, {4, 0, 0, 6, 0},
                   {4, -2, 1, -2, 10}, {4, 2, 1, -2, 10}, {4, -4, 1, 1, 70}, {4, 4, 1, 1, 70}}},
    {OUT_IS2, 0, 6, {{2, -2, 1, -3, 3}, {2, 2, 2, 1, 3}, {4, -2, 0, 1, 0}, {4, 2, 0, -1, 0}, {4, -4, 1, -1, 7},
                     {4, 4, 1, 1, 7}}},
    {OUT_1, 0, 6, {{2, -1, 1, 3, 6}, {2, 1, 1, -3, 6}, {4, -1, 0, -3, 0}, {4, 1, 0, 3, 0}, {4, -3, 1, 1, 7},
                   {4, 3, 1, -1, 7}}},
    {OUT_1, 1, 5, {{0, 0, 0, 14, 0}, {2, 0, 1, -4, 5}, {4, 0, 0, 2, 0}, {4, -4, 1, -1, 70}, {4, 4, 1, -1, 70}}},
    {OUT_I, 0, 6, {{2, -1, 1, -1, 6}, {2, 1, 1, -1, 6}, {4, -1, 0, 1, 0}, {4, 1, 0, 1, 0}, {4, -3, 1, -1, 7},
                   {4, 3, 1, -1, 7}}},
    {OUT_1, 1, 7, {{0, 0, 0, 14, 0}, {2, 0, 1, 2, 5}, {2, -2, 1, 1, 30}, {2, 2, 1, 1, 30}, {4, 0, 0, -8, 0},
                   {4, -2, 1, 2, 10}, {4, 2, 1, 2, 10}}},
    {OUT_IS2, 0, 6, {{2, -2, 1, -3, 3}, {2, 2, 2, 1, 3}, {4, -2, 0, 1, 0}, {4, 2, 0, -1, 0}, {4, -4, 1, 1, 7},
                     {4, 4, 1, -1, 7}}},
    {OUT_1, 0, 6, {{2, -1, 1, 1, 6}, {2, 1, 1, -1, 6}, {4, -1, 0, -1, 0}, {4, 1, 0, 1, 0}, {4, -3, 1, -1, 7},
                   {4, 3, 1, 1, 7}}},
    {OUT_IS2, 0, 4, {{2, -2, 1, -1, 3}, {2, 2, 1, 1, 3}, {4, -2, 0, -2, 0}, {4, 2, 0, 2, 0}}},
    {OUT_1, 0, 4, {{2, -1, 1, 3, 6}, {2, 1, 1, -3, 6}, {4, -1, 0, 4, 0}, {4, 1, 0, -4, 0}}},
    {OUT_1, 1, 9, {{0, 0, 0, 42, 0}, {2, 0, 1, -12, 5}, {2, -2, 1, -6, 30}, {2, 2, 1, -6, 30}, {4, 0, 0, 6, 0},
                   {4, -2, 1, 2, 10}, {4, 2, 1, 2, 10}, {4, -4, 1, 1, 70}, {4, 4, 1, 1, 70}}},
    {OUT_I, 0, 6, {{2, -1, 1, -3, 6}, {2, 1, 1, -3, 6}, {4, -1, 0, 3, 0}, {4, 1, 0, 3, 0}, {4, -3, 1, 1, 7},
                   {4, 3, 1, 1, 7}}},
    {OUT_1, 1, 7, {{0, 0, 0, 14, 0}, {2, 0, 1, 2, 5}, {2, -2, 1, -1, 30}, {2, 2, 1, -1, 30}, {4, 0, 0, -8, 0},
                   {4, -2, 1, -2, 10}, {4, 2, 1, -2, 10}}},
    {OUT_I, 0, 4, {{2, -1, 1, -3, 6}, {2, 1, 1, -3, 6}, {4, -1, 0, -4, 0}, {4, 1, 0, -4, 0}}},
    {OUT_1, 1, 3, {{0, 0, 0, 42, 0}, {2, 0, 1, 24, 5}, {4, 0, 0, 16, 0}}},
};

double term_coef(const A4Term& t) {
    float c = t.a;
    if (t.kind == 1) c = t.a * std::sqrt(t.b);
    else if (t.kind == 2) c = t.a * std::pow(t.b, 1.5f);
    return (double)c;
}

// complex l <= 4 coefficients (specfab order, 15) of real basis vector q
void real_unit_to_complex(int q, cd z[15]) {
    for (int j = 0; j < 15; j++) z[j] = 0;
    int l, am, c;
    real4_index(q, &l, &am, &c);
    const int ctr = (l == 0) ? 0 : (l == 2 ? 3 : 10);
    if (am == 0) { z[ctr] = 1.0; return; }
    cd v = c == 0 ? cd(1, 0) : cd(0, 1);
    z[ctr + am] = v;
    z[ctr - am] = ((am % 2) ? -1.0 : 1.0) * std::conj(v);
}

double fact(int n) {
    double f = 1;
    for (int i = 2; i <= n; i++) f *= i;
    return f;
}

// sqrt((l+m)!(l-m)!(l+n)!(l-n)!) with the square-free factor s entering as
// the reference's float32 sqrt (src/include/Dlmn_L4.f90 writes Sqrt((7.0)),
// Sqrt(17.5) = sqrtf(70)/2, (2.0)**1.5 = 2 sqrtf(2), ...); at l = 4,
// {|m|,|n|} = {2,1} the body divides by Sqrt((2.0)) instead, i.e. sqrt(2) is
// represented as 2/sqrtf(2).
double dnorm(int l, int m, int n) {
    long long P = (long long)std::llround(fact(l + m) * fact(l - m) * fact(l + n) * fact(l - n));
    long long sq = 1, root = 1, rem = P;
    for (long long p = 2; p * p <= rem; p++) {
        int e = 0;
        while (rem % p == 0) { rem /= p; e++; }
        for (int k = 0; k < e / 2; k++) root *= p;
        if (e & 1) sq *= p;
    }
    sq *= rem;
    if (sq == 1) return (double)root;
    const int am = std::abs(m), an = std::abs(n);
    if (l == 4 && am + an == 3 && am * an == 2) return root * (2.0 / (double)std::sqrt(2.0f));
    return root * (double)std::sqrt((float)sq);
}

// Wigner small-d d^l_{mn}(beta), Wigner's sum formula, normalised by dnorm
double small_d(int l, int m, int n, double beta) {
    const double c = std::cos(beta / 2), s = std::sin(beta / 2);
    double acc = 0;
    for (int k = std::max(0, n - m); k <= std::min(l + n, l - m); k++) {
        double term = std::pow(c, 2 * l + n - m - 2 * k) * std::pow(s, m - n + 2 * k) /
                      (fact(l + n - k) * fact(k) * fact(m - n + k) * fact(l - m - k));
        acc += ((m - n + k) & 1) ? -term : term;
    }
    return dnorm(l, m, n) * acc;
}

}  // namespace

void real4_index(int q, int* l, int* am, int* c) {
    if (q == 0) { *l = 0; *am = 0; *c = 0; return; }
    int base = (q <= 5) ? 1 : 6;
    *l = (q <= 5) ? 2 : 4;
    int r = q - base;
    *am = (r + 1) / 2;
    *c = (r == 0) ? 0 : ((r - 1) % 2);
}

// a2 tensor entries (xx, yy, zz, yz, xz, xy) of a state x are
// (A2T x) / x0 plus I/3 on the diagonal
void a2_tensor_map(double A2T[6][15]) {
    const double r2i = 0.70710678118654752440;
    for (int q = 0; q < 15; q++) {
        cd z[15];
        real_unit_to_complex(q, z);
        double v[6];
        a2_linear(z, v);
        for (int i = 0; i < 6; i++) A2T[i][q] = q == 0 ? 0.0 : (i < 3 ? v[i] : v[i] * r2i);
    }
}

// A4F[e][q]: index-sorted entry e of f_ev_c4 (src/moments.f90:202-209) for
// real basis vector q, times k / sqrt(4 pi) with k = Sqrt(Pi/5.)/21., so that
// the entries of a state x are (A4F x) / x0.
void a4full_map(double A4F[15][15]) {
    const double PI = 3.141592653589793;
    const double k = std::sqrt(PI / 5.0) / 21.0, c0 = std::sqrt(4 * PI);
    const double s2 = (double)std::sqrt(2.0f), s5 = (double)std::sqrt(5.0f);
    for (int q = 0; q < 15; q++) {
        cd z[15];
        real_unit_to_complex(q, z);
        for (int e = 0; e < 15; e++) {
            cd acc = 0;
            for (int t = 0; t < kA4[e].nterm; t++) {
                const A4Term& tm = kA4[e].t[t];
                const int ctr = (tm.l == 0) ? 0 : (tm.l == 2 ? 3 : 10);
                acc += term_coef(tm) * z[ctr + tm.m];
            }
            if (kA4[e].outer == OUT_I) acc = cd(0, 1) * acc;
            else if (kA4[e].outer == OUT_IS2) acc = cd(0, s2) * acc;
            if (kA4[e].div5) acc /= s5;
            A4F[e][q] = acc.real() * k / c0;
        }
    }
}

namespace {
// exponent of the prime p in k! (Legendre's formula)
int fact_exp(int k, int p) {
    int e = 0;
    for (long q = p; q <= k; q *= p) e += (int)(k / q);
    return e;
}

// dnorm for any l <= 15: sqrt((l+m)!(l-m)!(l+n)!(l-n)!) = root * sqrt(sq)
// from the prime exponents of the factorials (the product itself exceeds 64
// bits from l = 6 on), the square-free part sq as a float32 square root as
// in dnorm
long double dnorm_big(int l, int m, int n) {
    static const int primes[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29};
    long double root = 1;
    long long sq = 1;
    for (int p : primes) {
        const int e = fact_exp(l + m, p) + fact_exp(l - m, p) + fact_exp(l + n, p) + fact_exp(l - n, p);
        for (int i = 0; i < e / 2; i++) root *= p;
        if (e & 1) sq *= p;
    }
    if (sq == 1) return root;
    const int am = std::abs(m), an = std::abs(n);
    if (l == 4 && am + an == 3 && am * an == 2) return root * (2.0L / (long double)std::sqrt(2.0f));
    return root * (long double)std::sqrt((float)sq);
}
}  // namespace

void wigner_terms(int L, std::vector<WignerTerm>& terms, std::vector<int>& first) {
    terms.clear();
    first.assign(1, 0);
    for (int l = 2; l <= L; l += 2)
        for (int m = 0; m <= l; m++)
            for (int n = -l; n <= l; n++) {
                const long double nrm = dnorm_big(l, m, n);
                for (int k = std::max(0, n - m); k <= std::min(l + n, l - m); k++) {
                    long double den = 1;
                    for (int i = 2; i <= l + n - k; i++) den *= i;
                    for (int i = 2; i <= k; i++) den *= i;
                    for (int i = 2; i <= m - n + k; i++) den *= i;
                    for (int i = 2; i <= l - m - k; i++) den *= i;
                    const double c = (double)(nrm / den) * (((m - n + k) & 1) ? -1.0 : 1.0);
                    terms.push_back({c, 2 * l + n - m - 2 * k, m - n + 2 * k});
                }
                first.push_back((int)terms.size());
            }
}

// Real-basis rotation of the l = 2 and l = 4 blocks by rotate_nlm4(., theta,
// phi) (src/rotation.f90:14-27): n'_l^m = sum_n D^l(m,n) n_l^n with
// D^l(m,n) = exp(-i n phi) d^l_{mn}(theta).  Row/column order inside a block
// is the real basis order [Re n_l0, Re n_l1, Im n_l1, ...].
void rotation_l4(double theta, double phi, double R2[5][5], double R4[9][9]) {
    for (int q = 1; q < 15; q++) {
        cd z[15], w[15];
        real_unit_to_complex(q, z);
        for (int j = 0; j < 15; j++) w[j] = 0;
        for (int l = 2; l <= 4; l += 2) {
            const int ctr = (l == 2) ? 3 : 10;
            for (int m = -l; m <= l; m++) {
                cd acc = 0;
                for (int n = -l; n <= l; n++) {
                    if (z[ctr + n] == cd(0, 0)) continue;
                    cd ph = std::exp(cd(0, -(double)n * phi));
                    acc += ph * small_d(l, m, n, theta) * z[ctr + n];
                }
                w[ctr + m] = acc;
            }
        }
        // back to the real basis: entries m >= 0 of the rotated vector
        for (int p = 1; p < 15; p++) {
            int l, am, c;
            real4_index(p, &l, &am, &c);
            const int ctr = (l == 2) ? 3 : 10;
            double v = c == 0 ? w[ctr + am].real() : w[ctr + am].imag();
            int lq, amq, cq;
            real4_index(q, &lq, &amq, &cq);
            if (l != lq) continue;
            if (l == 2) R2[p - 1][q - 1] = v;
            else R4[p - 6][q - 6] = v;
        }
    }
}

}  // namespace sfg

namespace sfg {

// elas_fwdparams_tranisotropic / elas_revparams_tranisotropic
// (src/elasticities.f90:64-96) from Lame = (lam, mu, Elam, Emu, Egam).
void elas_params(const double Lame[5], double kf[5], double kr[5]) {
    const double lam = Lame[0], mu = Lame[1], El = Lame[2], Em = Lame[3], Eg = Lame[4];
    const double gam = lam + 2 * mu;
    const double xi = 2 * Eg * gam * (lam + mu) - 2 * (El * lam) * (El * lam);
    kf[0] = ((El * lam) * (El * lam) - Eg * gam * lam) / (2 * mu * xi);
    kf[1] = 1 / (2 * mu);
    kf[2] = (Eg * gam * lam - El * lam * (El * lam + 2 * mu)) / (2 * mu * xi);
    kf[3] = (4 * mu * lam * (1 + El) + 4 * mu * mu + Eg * gam * gam - El * El * lam * lam) / (2 * mu * xi) -
            1 / (Em * mu);
    kf[4] = (1 - Em) / (2 * Em * mu);
    kr[0] = lam;
    kr[1] = 2 * mu;
    kr[2] = -lam * (1 - El);
    kr[3] = (1 + Eg) * gam - 2 * (El * lam + 2 * Em * mu);
    kr[4] = -2 * mu * (1 - Em);
}

}  // namespace sfg
