// Internal declarations shared by the host basis builder, the kernels and the
// C-ABI layer of libspecfab_gpu.
#pragma once

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace sfg {

constexpr int LMAX_TABLE = 20;   // coupling tables are generated up to L = 20
constexpr int NMAX = 231;        // nlm_len(20)
constexpr int NPAR_LROT = 8;     // real parameters of the lattice-rotation operator
constexpr int NPAR_DDRX = 15;    // real parameters of the DDRX source operator
constexpr int NPAR = NPAR_LROT + NPAR_DDRX;
constexpr int NOMEGA = 15;       // weights of <D> over the l <= 4 real coefficients

inline int nlm_len(int L) { return (L + 1) * (L + 2) / 2; }

// One stored coupling-table entry (table 0..3 = GC, GCm, GC_m1, GC_p1).
struct GauntEntry { int t, i, j, k; double v; };
const std::vector<GauntEntry>& gaunt_entries();

// Complex spectral index <-> (l, m), specfab order (src/header.f90:30).
struct LM { int l, m; };
std::vector<LM> lm_list(int L);

// Real ("reality-reduced") basis: for each even l the coefficients
//   x = [Re n_l0, Re n_l1, Im n_l1, ..., Re n_ll, Im n_ll]
// hold the n real degrees of freedom of a real ODF, n_l^-m = (-1)^m conj(n_l^m).
struct RealIndex { int l, m, c; };   // c = 0 real part, 1 imaginary part

// Operator of one truncation L in the real basis:
//   d x_a / dt = sum_p theta_p sum_b R_p[a][b] x_b  +  s_{l(a)} x_a
// p = 0..7 lattice rotation (real parts of the quadric coefficients),
// p = 8..22 DDRX source (Gamma0-scaled harmonic weights).  Terms are stored
// grouped by output row a, then by parameter p.
struct Term { int col; double val; };
struct Group { int param; int t0, t1; };       // terms [t0, t1)
struct Row { int g0, g1; int l; };             // groups [g0, g1)

struct RealOperator {
    int L = 0, n = 0;
    std::vector<RealIndex> ridx;                // size n
    std::vector<Row> rows;                      // size n
    std::vector<Group> groups;
    std::vector<Term> terms;
    int nterms_lrot = 0, nterms_ddrx = 0;
    // complex index j -> (real index of Re, real index of Im or -1, sign, conj)
    std::vector<int> cre, cim;                  // size n (complex count)
    std::vector<int> csgn;                      // (-1)^m for m < 0 else 1
    std::vector<int> cneg;                      // 1 if m < 0 (conjugate)
    // <D> linear maps: a2 Mandel (6 x 15) and a4 Mandel (36 x 15) over the
    // first 15 real coefficients, already divided by the normalisation so
    // that a2v = A2 x / x0 + I/3 and a4v = A4 x / x0.
    double A2[6][15];
    double A4[36][15];
    // REG shape per l: |l(l+1)/(L(L+1))|^expo (src/dynamics.f90:510-513)
    std::vector<double> reg_shape;              // size L/2+1
    double reg_nu = 0, reg_expo = 0;
    bool reg_defined = false;
};

const RealOperator& real_operator(int L);  // built once per L, thread-safe

// Real l <= 4 block (15 DOFs, real basis order): q -> (l, |m|, re/im part).
void real4_index(int q, int* l, int* am, int* c);
// Full fourth-order structure tensor (src/moments.f90:202-209, REAL(4)
// literals of src/include/ev_c4__body.f90): index-sorted entries
// (1111,1112,1113,1122,1123,1133,1222,1223,1233,1333,2222,2223,2233,2333,3333)
// of a state x are (A4F x) / x0.
void a4full_map(double A4F[15][15]);
// a2 body (src/include/ev_c2__body.f90) over the complex l <= 4 block:
// Mandel vector (xx, yy, zz, sqrt2 yz, sqrt2 xz, sqrt2 xy) without I/3 and 1/n00
void a2_linear(const std::complex<double>* z, double v[6]);
// a2 tensor entries (xx, yy, zz, yz, xz, xy) of a state x are (A2T x) / x0 + I/3
void a2_tensor_map(double A2T[6][15]);
// rotate_nlm4 (src/rotation.f90:14-27) as real 5x5 / 9x9 blocks.
void rotation_l4(double theta, double phi, double R2[5][5], double R4[9][9]);
// Terms of the Wigner small-d entries d^l_{mn}(theta), l = 2, 4, ..., L,
// m = 0..l, n = -l..l (entry index in that order): d = sum of
// coef * cos(theta/2)^a * sin(theta/2)^b over the entry's terms (Wigner's sum
// formula; the normalisation with the reference's REAL(4) square roots where
// src/include/Dlmn.f90 has them as single factors).  first[e] .. first[e+1]
// are entry e's terms.
struct WignerTerm { double coef; int a, b; };
void wigner_terms(int L, std::vector<WignerTerm>& terms, std::vector<int>& first);
// Reuss (forward) and Voigt (reverse) elastic coefficients k1..k5.
void elas_params(const double Lame[5], double kf[5], double kr[5]);

// LROT harmonic weights (g0, gz, gn, gp)(k = 0..5) for given quadric
// coefficients qe(-2..2), qo(-1..1); src/dynamics.f90:78-91.
void lrot_weights(const std::complex<double> qe[5], const std::complex<double> qo[3],
                  std::complex<double> g0[6], std::complex<double> gz[6],
                  std::complex<double> gn[6], std::complex<double> gp[6]);

}  // namespace sfg
