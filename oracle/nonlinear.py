"""Numpy restatement of the nonlinear-grain homogenization path: n' = 3 Sachs
(src/homogenizations.f90:69-116, the __isotropic normalisation :191-222), the
basal-RSS deformabilities <D> (src/dynamics.f90:385-449) and E_CAFFE
(src/enhancementfactors.f90:302-332).

Test infrastructure: imported only by tests/ as the checker; nothing under
specfab_amd/ imports it.

  moments     the exact k-th moments <c^k> (k = 2..8) of an ODF nlm, by
              quadrature of psi = sum n_lm Y_lm (scipy's sph_harm_y; Gauss-
              Legendre in cos(theta) x trapezoid in phi, exact for the
              polynomial integrands): independent of the tables that
              specfab_amd/tables/make_moment_maps.py writes.
  a2, a4      the reference's structure-tensor bodies (f_ev_c2, f_ev_c4
              with its REAL(4) literals and its a4(3,2,1,2) entry), from the C
              oracle (oracle/sf_oracle.c, pinned there against oracle/_ref).
  sachs3      eps(tau) of the n' = 3 Sachs average, written as the reference
              contracts it (doubleinner42 / 62 / 82 and doubleinner22 with
              their index order, src/tensorproducts.f90:153-260).
  Eij_sachs3  Evw = (eps : vw) / (eps_iso : vw) for the six stress states of a
              frame (src/enhancementfactors.f90:47-69, 398-413), alpha = 0.
  ev_D        n_RSS = 2: the C oracle's ev_D2; n_RSS = 4: ev_D4 from the body
              a4 and the exact a6 / a8.
  E_CAFFE     the Placidi et al. (2010) map of <D>.
The reference's a6 / a8 bodies (src/include/ev_c6__body.f90, ev_c8__body.f90)
carry REAL(4) literal coefficients; these exact moments differ from them at
the 1e-7 level, which is the parity bar of everything that uses them.
"""
import numpy as np
from scipy.special import sph_harm_y

NG, NP = 24, 48


def _grid():
    xg, wg = np.polynomial.legendre.leggauss(NG)
    ph = np.arange(NP) * 2 * np.pi / NP
    TH = np.repeat(np.arccos(xg)[:, None], NP, axis=1)
    PH = np.repeat(ph[None, :], NG, axis=0)
    W = wg[:, None] * np.full(NP, 2 * np.pi / NP)[None, :]
    C = np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)])
    return TH, PH, W, C


_G = None


def moments(nlm, ks=(2, 4, 6, 8)):
    """{k: <c^k> as a (3,)*k array} of the ODF with coefficients nlm (any even L)."""
    global _G
    if _G is None:
        _G = _grid()
    TH, PH, W, C = _G
    nlm = np.asarray(nlm)
    L = int(round((-3 + np.sqrt(1 + 8 * len(nlm))) / 2))
    lm = [(l, m) for l in range(0, L + 1, 2) for m in range(-l, l + 1)]
    psi = np.real(sum(z * sph_harm_y(l, m, TH, PH) for z, (l, m) in zip(nlm, lm))) * W
    psi = psi / psi.sum()
    out = {}
    for k in ks:
        letters = "abcdefgh"[:k]
        out[k] = np.einsum(",".join(c + "xy" for c in letters) + ",xy->" + letters, *([C] * k), psi)
    return out


def rheo_params(Eg, n, ef):
    """rheo_params_tranisotropic, src/rheologies.f90:123-135 (d = 3)."""
    ne = ef * 2 / (n + 1)
    return (Eg[0] ** ne - 1) / 2, (3 * (Eg[0] ** ne + 1) - 2) / 2 - 2 * Eg[1] ** ne, Eg[1] ** ne - 1


def _d22(A, B):          # A_ij B_ji
    return np.einsum("ij,ji->", A, B)


def _d42(A, B):          # A_lkij B_ji
    return np.einsum("lkij,ji->lk", A, B)


def _d62(A, B):
    return np.einsum("nmlkij,ji->nmlk", A, B)


def _d82(A, B):
    return np.einsum("ponmlkij,ji->ponmlk", A, B)


def sachs3(tau, a2, a4, a6, a8, Eg):
    """eps(tau) of the n' = 3 Sachs average, src/homogenizations.f90:93-103,115."""
    A, B, C = rheo_params(Eg, 3.0, 1)
    I2 = _d22(tau, tau)
    t2 = tau @ tau
    c0 = I2 * 1.0 + B * _d22(_d42(a4, tau), tau) + 2 * C * _d22(a2, t2)
    c2 = I2 * a2 + B * _d42(_d62(a6, tau), tau) + 2 * C * _d42(a4, t2)
    c4 = I2 * a4 + B * _d62(_d82(a8, tau), tau) + 2 * C * _d62(a6, t2)
    a4tau = _d42(c4, tau)
    return c0 * tau - A * _d22(c2, tau) * np.eye(3) + B * a4tau + C * (tau @ c2 + c2 @ tau)


def _states(e1, e2, e3):
    e = [np.asarray(e1, float), np.asarray(e2, float), np.asarray(e3, float)]
    for (iv, iw) in [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]:
        v, w = e[iv], e[iw]
        tau = np.eye(3) / 3 - np.outer(v, v) if iv == iw else np.outer(v, w) + np.outer(w, v)
        yield tau, np.outer(v, w)


def _structure(nlm, oracle):
    m = moments(nlm, (6, 8))
    return oracle.a2(nlm), oracle.a4(nlm), m[6], m[8]


def isotropic(L):
    z = np.zeros((L + 1) * (L + 2) // 2, dtype=complex)
    z[0] = 1 / np.sqrt(4 * np.pi)
    return z


def Eij_sachs3(nlm, e1, e2, e3, Eg, oracle):
    """Eij (6,) of n' = 3 grains, Sachs (alpha = 0), src/enhancementfactors.f90:23-69."""
    L = int(round((-3 + np.sqrt(1 + 8 * len(nlm))) / 2))
    S = _structure(nlm, oracle)
    Si = _structure(isotropic(L), oracle)
    return np.array([_d22(sachs3(tau, *S, Eg), vw) / _d22(sachs3(tau, *Si, Eg), vw)
                     for tau, vw in _states(e1, e2, e3)])


def ev_D(nlm, tau, n_rss, oracle):
    """ev_D(nlm, tau, n_RSS), src/dynamics.f90:385-449."""
    if n_rss == 2:
        return oracle.ev_D2(nlm, tau)
    a2, a4, a6, a8 = _structure(nlm, oracle)
    t2 = tau @ tau
    norm = np.trace(t2)
    tt = np.einsum("ij,kl->ijkl", tau, tau)
    t2t = np.einsum("ij,kl->ijkl", t2, tau)
    D = _d22(t2, _d42(a4, t2))
    D = D + np.einsum("lkij,jikl->", tt, np.einsum("ponmlkij,jikl->ponm", a8, tt))
    D = D - 2 * np.einsum("lkij,jikl->", t2t, _d62(a6, tau))
    return 35 / 2.0 * D / norm ** 2


def E_CAFFE(nlm, eps, Emin, Emax, n_grain, oracle):
    """E_CAFFE (src/enhancementfactors.f90:302-332)."""
    n_rss = 4 if n_grain == 3 else 2
    D = ev_D(nlm, eps, n_rss, oracle)
    Dmaxpow = {2: 5 / 2.0, 4: 35 / 8.0}[n_rss] ** (4.0 / n_rss)
    if D < 1:
        gam = (4.0 / n_rss) / Dmaxpow * (Emax - 1) / (1 - Emin)
        return Emin + (1 - Emin) * D ** gam
    return ((Emax - 1) * D ** (4.0 / n_rss) + Dmaxpow - Emax) / (Dmaxpow - 1)
