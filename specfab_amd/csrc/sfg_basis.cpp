// Host-side construction of the real-basis fabric operators from the
// spectral coupling tables.  Runs once per truncation L (cached).
//
// The reference assembles complex n x n operators per parcel,
//   M_LROT  = -sum_k GC(:,:,k) g0_k + GCm gz_k + GC_m1 gn_k + GC_p1 gp_k
//             (src/dynamics.f90:94-96, weights :78-91)
//   M_DDRX  =  sum_k GC(:,:,k) g_k            (src/dynamics.f90:295-297)
// with complex weights that are linear in the quadric coefficients of the
// strain rate / spin / stress.  Because the ODF is real, n_l^-m =
// (-1)^m conj(n_l^m), and the weights obey the same reality relation, both
// operators are exactly linear maps of the n real degrees of freedom with
// real coefficients.  We therefore precompute, per L, the real tensors
// R_p[a][b] (p = 8 LROT + 15 DDRX real parameters) and let the GPU evaluate
//   f_a = sum_p theta_p (sum_b R_p[a][b] x_b)
// with theta_p computed per parcel from the reference's weight formulas.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <mutex>
#include <stdexcept>

#include "sfg_internal.h"

namespace sfg {

namespace {
struct RawEntry { int t, i, j, k; uint32_t bits; };
const RawEntry kGaunt[] = {
#include "generated/gaunt_L20.inc"
};
}  // namespace

const std::vector<GauntEntry>& gaunt_entries() {
    static std::vector<GauntEntry> v = [] {
        std::vector<GauntEntry> out;
        out.reserve(sizeof(kGaunt) / sizeof(kGaunt[0]));
        for (const auto& e : kGaunt) {
            float f;
            std::memcpy(&f, &e.bits, 4);
            out.push_back({e.t, e.i, e.j, e.k, (double)f});   // REAL(4) literal widened
        }
        return out;
    }();
    return v;
}

std::vector<LM> lm_list(int L) {
    std::vector<LM> v;
    for (int l = 0; l <= L; l += 2)
        for (int m = -l; m <= l; m++) v.push_back({l, m});
    return v;
}

using cd = std::complex<double>;

void lrot_weights(const cd qe[5], const cd qo[3], cd g0[6], cd gz[6], cd gn[6], cd gp[6]) {
    // REAL(4) literals of src/dynamics.f90:79-86
    const double s3 = (double)std::sqrt(3.0f), s6 = (double)std::sqrt(6.0f);
    const double c56 = (double)std::sqrt(5.0f / 6.0f), c23 = (double)std::sqrt(2.0f / 3.0f);
    const double c32 = (double)std::sqrt(3.0f / 2.0f);
    const cd I(0, 1);
    cd gz0 = -((I * s3) * qo[1]);
    cd gn0 = -(((I * 6.0) / s6) * qo[0]);
    cd gp0 = ((I * 6.0) / s6) * qo[2];
    g0[0] = 0;
    for (int k = 1; k < 6; k++) g0[k] = 3.0 * qe[k - 1];
    gz[0] = gz0; gz[1] = -qe[0]; gz[2] = 0; gz[3] = 0; gz[4] = 0; gz[5] = qe[4];
    gn[0] = gn0 + c56 * qe[1]; gn[1] = 0; gn[2] = qe[0]; gn[3] = c23 * qe[1]; gn[4] = c32 * qe[2]; gn[5] = 2.0 * qe[3];
    gp[0] = gp0 + c56 * qe[3]; gp[1] = 2.0 * qe[1]; gp[2] = c32 * qe[2]; gp[3] = c23 * qe[3]; gp[4] = qe[4]; gp[5] = 0;
}

// Linear part of the Mandel a2 body over the complex l <= 4 block
// (src/include/ev_c2__body.f90 without the I/3 term and the 1/n00):
// v = (xx, yy, zz, sqrt2 yz, sqrt2 xz, sqrt2 xy).
void a2_linear(const cd* z, double v[6]) {
    const double c = std::sqrt(2.0 / 3), s = std::sqrt(2 / 15.0), r2 = std::sqrt(2.0);
    cd h0 = z[3], h1 = z[4], h2 = z[5];
    double e11 = -0.5 * c * h0.real() + h2.real();
    double e22 = -0.5 * c * h0.real() - h2.real();
    double e33 = c * h0.real();
    double e12 = -h2.imag(), e13 = -h1.real(), e23 = h1.imag();
    v[0] = s * e11; v[1] = s * e22; v[2] = s * e33;
    v[3] = r2 * (s * e23); v[4] = r2 * (s * e13); v[5] = r2 * (s * e12);
}

namespace {

int regcalib(int L, double* expo, double* nu) {
    // src/include/regcalib.f90:1-36
    switch (L) {
    case 4: *expo = 1.700; *nu = 1.9879322126397958; return 0;
    case 6: *expo = 1.150; *nu = 3.0011508426238862; return 0;
    case 8: *expo = 1.600; *nu = 5.7498069921352384; return 0;
    case 10: *expo = 2.000; *nu = 10.7048905312159288; return 0;
    case 12: *expo = 2.000; *nu = 10.6068117205577668; return 0;
    case 14: *expo = 2.000; *nu = 13.3591023418822363; return 0;
    case 16: *expo = 2.500; *nu = 15.3094482670021108; return 0;
    case 18: *expo = 2.500; *nu = 16.4844589176829217; return 0;
    case 20: *expo = 3.000; *nu = 19.9467342880730136; return 0;
    default: return -1;
    }
}

// Linear part of the Mandel a4 body over the complex l <= 4 block:
// src/include/ev_c4_Mandel__body.f90 times k/sqrt(4 pi) (src/moments.f90:211-218).
void a4_linear(const cd* z, double a4v[36]) {
    const double PI = 3.141592653589793;
    double n00 = z[0].real();
    auto RE = [&](int m) { return z[3 + m].real(); };
    auto IM = [&](int m) { return z[3 + m].imag(); };
    auto R4 = [&](int m) { return z[10 + m].real(); };
    auto I4 = [&](int m) { return z[10 + m].imag(); };
    const double s5 = std::sqrt(5.0), s6 = std::sqrt(6.0), s7 = std::sqrt(7.0), s10 = std::sqrt(10.0);
    const double s30 = std::sqrt(30.0), s70 = std::sqrt(70.0), s3 = std::sqrt(3.0), s15 = std::sqrt(15.0);
    double e[6][6];
    e[0][0] = 21.0 * n00 + s5 * (-6.0) * RE(0) + s30 * 6.0 * RE(2) + 3.0 * R4(0) + s10 * (-2.0) * R4(2) + s70 * R4(4);
    e[0][1] = 7.0 * n00 + (-2.0) * s5 * RE(0) + R4(0) + (-1.0) * s70 * R4(4);
    e[0][2] = 7.0 * n00 + s5 * (RE(0) + s6 * RE(2)) + (-4.0) * R4(0) + s10 * 2.0 * R4(2);
    e[0][3] = s10 * (s6 * IM(1) + (-1.0) * I4(1) + s7 * I4(3));
    e[0][4] = s10 * (-1.0) * (3.0 * s6 * RE(1) + (-3.0) * R4(1) + s7 * R4(3));
    e[0][5] = (-2.0) * s5 * (std::pow(3.0, 1.5) * IM(2) + (-1.0) * I4(2) + s7 * I4(4));
    e[1][1] = 21.0 * n00 + s5 * (-6.0) * (RE(0) + s6 * RE(2)) + 3.0 * R4(0) + s10 * (2.0 * R4(2) + s7 * R4(4));
    e[1][2] = 7.0 * n00 + s5 * (RE(0) + (-1.0) * s6 * RE(2)) + (-2.0) * (2.0 * R4(0) + s10 * R4(2));
    e[1][3] = s10 * (3.0 * s6 * IM(1) + (-3.0) * I4(1) + (-1.0) * s7 * I4(3));
    e[1][4] = s15 * (-2.0) * RE(1) + s10 * (R4(1) + s7 * R4(3));
    e[1][5] = 2.0 * s5 * ((-3.0) * s3 * IM(2) + I4(2) + s7 * I4(4));
    e[2][2] = 21.0 * n00 + 12.0 * s5 * RE(0) + 8.0 * R4(0);
    e[2][3] = s10 * (3.0 * s6 * IM(1) + 4.0 * I4(1));
    e[2][4] = s10 * (-1.0) * (3.0 * s6 * RE(1) + 4.0 * R4(1));
    e[2][5] = (-2.0) * s5 * (s3 * IM(2) + 2.0 * I4(2));
    e[3][3] = 2.0 * (7.0 * n00 + s5 * (RE(0) + (-1.0) * s6 * RE(2)) + (-2.0) * (2.0 * R4(0) + s10 * R4(2)));
    e[3][4] = s10 * (-2.0) * (s3 * IM(2) + 2.0 * I4(2));
    e[3][5] = s5 * ((-2.0) * s6 * RE(1) + 2.0 * (R4(1) + s7 * R4(3)));
    e[4][4] = 2.0 * (7.0 * n00 + s5 * (RE(0) + s6 * RE(2)) + (-4.0) * R4(0) + s10 * 2.0 * R4(2));
    e[4][5] = 2.0 * s5 * (s6 * IM(1) + (-1.0) * I4(1) + s7 * I4(3));
    e[5][5] = 2.0 * (7.0 * n00 + (-2.0) * s5 * RE(0) + R4(0) + (-1.0) * s70 * R4(4));
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < i; j++) e[i][j] = e[j][i];
    const double k = (2 * std::sqrt(PI)) / 105.0;
    const double c0 = std::sqrt(4 * PI);
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) a4v[6 * i + j] = e[i][j] * k / c0;
}

RealOperator build(int L) {
    if (L < 0 || L > LMAX_TABLE || (L % 2)) throw std::invalid_argument("L must be even, 0 <= L <= 20");
    RealOperator op;
    op.L = L;
    op.n = nlm_len(L);
    const int n = op.n;
    auto lm = lm_list(L);
    // real index order
    std::map<std::tuple<int, int, int>, int> rpos;
    for (int l = 0; l <= L; l += 2)
        for (int m = 0; m <= l; m++) {
            rpos[{l, m, 0}] = (int)op.ridx.size();
            op.ridx.push_back({l, m, 0});
            if (m > 0) {
                rpos[{l, m, 1}] = (int)op.ridx.size();
                op.ridx.push_back({l, m, 1});
            }
        }
    // complex -> real map E: n_j = sum_b E[j][b] x_b
    std::vector<std::vector<std::pair<int, cd>>> E(n);
    op.cre.assign(n, -1); op.cim.assign(n, -1); op.csgn.assign(n, 1); op.cneg.assign(n, 0);
    for (int j = 0; j < n; j++) {
        int l = lm[j].l, m = lm[j].m, am = std::abs(m);
        if (am == 0) {
            E[j].push_back({rpos[{l, 0, 0}], cd(1, 0)});
            op.cre[j] = rpos[{l, 0, 0}];
        } else {
            double s = (m < 0 && (am % 2)) ? -1.0 : 1.0;
            E[j].push_back({rpos[{l, am, 0}], cd(s, 0)});
            E[j].push_back({rpos[{l, am, 1}], cd(0, m > 0 ? s : -s)});
            op.cre[j] = rpos[{l, am, 0}];
            op.cim[j] = rpos[{l, am, 1}];
            op.csgn[j] = (int)s;
            op.cneg[j] = m < 0;
        }
    }
    // parameter -> complex weights per (table, k): d[p][t][k]
    static cd d[NPAR][4][15];
    for (auto& a : d) for (auto& b : a) for (auto& c : b) c = 0;
    for (int p = 0; p < NPAR_LROT; p++) {
        cd qe[5] = {0, 0, 0, 0, 0}, qo[3] = {0, 0, 0};
        const cd I(0, 1);
        switch (p) {
        case 0: qe[2] = 1; break;
        case 1: qe[3] = 1; qe[1] = -1; break;            // Re qe(+1); qe(-1) = -conj(qe(+1))
        case 2: qe[3] = I; qe[1] = I; break;             // Im qe(+1)
        case 3: qe[4] = 1; qe[0] = 1; break;             // Re qe(+2); qe(-2) = conj(qe(+2))
        case 4: qe[4] = I; qe[0] = -I; break;            // Im qe(+2)
        case 5: qo[1] = 1; break;
        case 6: qo[2] = 1; qo[0] = -1; break;            // Re qo(+1); qo(-1) = -conj(qo(+1))
        case 7: qo[2] = I; qo[0] = I; break;             // Im qo(+1)
        }
        cd g[4][6];
        lrot_weights(qe, qo, g[0], g[1], g[2], g[3]);
        for (int t = 0; t < 4; t++)
            for (int k = 0; k < 6; k++) d[p][t][k] = -g[t][k];   // leading minus of M_LROT
    }
    {
        // DDRX: theta = (Re g(l,m), Im g(l,m)) for m >= 0, g(l,-m) = (-1)^m conj g(l,m)
        int p = NPAR_LROT;
        auto kidx = [](int l, int m) { return l == 0 ? 0 : (l == 2 ? 1 + m + 2 : 6 + m + 4); };
        for (int l = 0; l <= 4; l += 2)
            for (int m = 0; m <= l; m++) {
                double s = (m % 2) ? -1.0 : 1.0;
                if (m == 0) {
                    d[p++][0][kidx(l, 0)] = 1;
                } else {
                    d[p][0][kidx(l, m)] = 1; d[p][0][kidx(l, -m)] = s; p++;
                    d[p][0][kidx(l, m)] = cd(0, 1); d[p][0][kidx(l, -m)] = cd(0, -s); p++;
                }
            }
    }
    // accumulate R_p[a][b] (dense per p, then compress)
    std::vector<double> R((size_t)NPAR * n * n, 0.0);
    auto Rp = [&](int p, int a, int b) -> double& { return R[((size_t)p * n + a) * n + b]; };
    // rows a for complex i with m_i >= 0
    std::vector<int> row_re(n, -1), row_im(n, -1);
    for (int i = 0; i < n; i++)
        if (lm[i].m >= 0) {
            row_re[i] = rpos[{lm[i].l, lm[i].m, 0}];
            if (lm[i].m > 0) row_im[i] = rpos[{lm[i].l, lm[i].m, 1}];
        }
    for (const auto& e : gaunt_entries()) {
        if (e.i >= n || e.j >= n || row_re[e.i] < 0) continue;
        for (int p = 0; p < NPAR; p++) {
            cd w = d[p][e.t][e.k];
            if (w == cd(0, 0)) continue;
            cd c = e.v * w;
            for (auto& be : E[e.j]) {
                cd ce = c * be.second;
                Rp(p, row_re[e.i], be.first) += ce.real();
                if (row_im[e.i] >= 0) Rp(p, row_im[e.i], be.first) += ce.imag();
            }
        }
    }
    // compress: rows -> groups (param) -> terms
    for (int a = 0; a < n; a++) {
        Row row;
        row.g0 = (int)op.groups.size();
        row.l = op.ridx[a].l;
        for (int p = 0; p < NPAR; p++) {
            Group g;
            g.param = p;
            g.t0 = (int)op.terms.size();
            for (int b = 0; b < n; b++) {
                double v = Rp(p, a, b);
                if (v != 0.0) {
                    op.terms.push_back({b, v});
                    (p < NPAR_LROT ? op.nterms_lrot : op.nterms_ddrx)++;
                }
            }
            g.t1 = (int)op.terms.size();
            if (g.t1 > g.t0) op.groups.push_back(g);
        }
        row.g1 = (int)op.groups.size();
        op.rows.push_back(row);
    }
    // <D> maps over the first 15 real coefficients (L >= 4), else over what exists
    for (auto& r : op.A2) for (auto& v : r) v = 0;
    for (auto& r : op.A4) for (auto& v : r) v = 0;
    const int nq = std::min(n, 15);
    for (int q = 0; q < nq; q++) {
        std::vector<double> xq(15, 0.0);
        xq[q] = 1.0;
        cd z[15];
        auto lm4 = lm_list(4);
        for (int j = 0; j < 15; j++) {
            cd v = 0;
            int l = lm4[j].l, m = lm4[j].m, am = std::abs(m);
            if (l > L) { z[j] = 0; continue; }
            int ire = rpos[{l, am, 0}];
            if (am == 0) v = xq[ire];
            else {
                int iim = rpos[{l, am, 1}];
                double s = (m < 0 && (am % 2)) ? -1.0 : 1.0;
                v = m > 0 ? cd(xq[ire], xq[iim]) : s * cd(xq[ire], -xq[iim]);
            }
            z[j] = v;
        }
        double a2v[6], a4v[36];
        a2_linear(z, a2v);
        a4_linear(z, a4v);
        for (int i = 0; i < 6; i++) op.A2[i][q] = (q == 0) ? 0.0 : a2v[i];
        for (int i = 0; i < 36; i++) op.A4[i][q] = a4v[i];
    }
    // REG
    op.reg_defined = regcalib(L, &op.reg_expo, &op.reg_nu) == 0;
    for (int l = 0; l <= L; l += 2) {
        double Ld = -(double)(l * (l + 1));
        op.reg_shape.push_back(op.reg_defined ? std::pow(std::fabs(Ld / (L * (L + 1))), op.reg_expo) : 0.0);
    }
    return op;
}

}  // namespace

const RealOperator& real_operator(int L) {
    static std::mutex mu;
    static std::map<int, RealOperator> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(L);
    if (it == cache.end()) it = cache.emplace(L, build(L)).first;
    return it->second;
}

}  // namespace sfg
