"""Numpy restatement of the orthotropic-grain (olivine) homogenization:
Eij_orthotropic (Sachs, src/enhancementfactors.f90:134-188 with
src/homogenizations.f90:288-319) and Vi_elastic_orthotropic (Voigt,
src/wavepropagation.f90:41-68, 98-115), and their discrete-grain forms
(src/enhancementfactors.f90:191-266, src/wavepropagation.f90:70-96,
src/moments.f90:386-417).

Test infrastructure: imported only by tests/ as the checker; nothing under
specfab_amd/ imports it.

  ai          ai_orthotropic (src/moments.f90:357-384): a2_i and a4_ii of each
              given distribution are the reference's a2 / a4 bodies (the C
              oracle's restatement: a4 with its REAL(4) literals and its
              a4(3,2,1,2) entry); the joint tensors a4_jk and, when q3 is not
              given, a2 / a4 of v = b x n follow from the two distributions'
              moments under the product density with the kernel sin^2 of the
              axes' angle:
                a4_joint(b, n)      = (a2b (x) a2n - a4b : a4n) / (1 - a2b : a2n)
                a4_jointcross(b, n) = eps eps (a4n (x) a2b) / (1 - a2b : a2n)   (<n n v v>)
                a2_orth(b, n)       = eps eps (a2b (x) a2n) / (1 - a2b : a2n)
                a4_orth(b, n)       = eps^4 (a4b (x) a4n) / (1 - 2 a2b : a2n + a4b :: a4n)
              (the content of the generated bodies of src/moments.f90:242-310,
              identified against oracle/_ref to their REAL(4) level; the
              symmetric a4 entering them is the body's, without its (3,2,1,2)
              entry).  The derived branch assigns a4_jk(1) = a4_jointcross(q2, q1)
              and a4_jk(2) = a4_jointcross(q1, q2) as the reference does.
  sachs_Q     Q = sum_i ci M2_i + ci+3 H2_i (rheo_params_orthotropic,
              src/rheologies.f90:186-206; a4_sym2 / a4_sym4,
              src/tensorproducts.f90:53-81).
  Eij         Evw = (doubleinner42(Q, tau) : vw) / (the same for the
              isotropic fabric), six stress states of the frame.
  Vi          rotate_nlm4 of each distribution (the C oracle's Wigner-D with
              the reference's REAL(4) rule), the Voigt acoustic tensor for
              k = z, and the reference's Cardano solution (qnorm_to_vi) in its
              (qS1, qS2, qP) order; its REAL(4) exponents move the roots by
              ~1e-8 relative from the exact eigenvalues, which
              qnorm_to_vi's test pins against numpy's eigvalsh.
"""
import numpy as np

from oracle import nonlinear as N

EPS = np.zeros((3, 3, 3))
EPS[0, 1, 2] = EPS[1, 2, 0] = EPS[2, 0, 1] = 1
EPS[0, 2, 1] = EPS[2, 1, 0] = EPS[1, 0, 2] = -1


def _sorted_sym(a4):
    """the symmetric tensor holding a4's index-sorted entries"""
    out = np.empty_like(a4)
    for idx in np.ndindex(3, 3, 3, 3):
        out[idx] = a4[tuple(sorted(idx))]
    return out


def dist(nlm, oracle):
    """(a2, a4 body, a4 symmetric) of one distribution (its l <= 4 block)"""
    z = np.zeros(15, dtype=complex)
    z[:min(15, len(nlm))] = np.asarray(nlm)[:15]
    a4 = oracle.a4(z)
    return oracle.a2(z), a4, _sorted_sym(a4)


def a4_joint(b, n):
    return (np.einsum("ij,kl->ijkl", b[0], n[0]) - np.einsum("ijpq,klpq->ijkl", b[2], n[2])) / \
        (1 - np.einsum("ij,ij", b[0], n[0]))


def a4_jointcross(b, n):
    return np.einsum("kpq,lrs,ijqs,pr->ijkl", EPS, EPS, n[2], b[0]) / (1 - np.einsum("ij,ij", b[0], n[0]))


def a2_orth(b, n):
    return np.einsum("ipq,jrs,pr,qs->ij", EPS, EPS, b[0], n[0]) / (1 - np.einsum("ij,ij", b[0], n[0]))


def a4_orth(b, n):
    n4 = 1 - 2 * np.einsum("ij,ij", b[0], n[0]) + np.einsum("ijkl,ijkl", b[2], n[2])
    return np.einsum("iab,jcd,kef,lgh,aceg,bdfh->ijkl", EPS, EPS, EPS, EPS, b[2], n[2]) / n4


def ai(q1, q2, q3, oracle):
    """(a2_i (3,3,3), a4_ii (3,3,3,3,3), a4_jk (3,3,3,3,3)) of ai_orthotropic"""
    d1, d2 = dist(q1, oracle), dist(q2, oracle)
    a2_i = np.zeros((3, 3, 3))
    a4_ii = np.zeros((3,) + (3,) * 4)
    a4_jk = np.zeros((3,) + (3,) * 4)
    a2_i[0], a2_i[1], a4_ii[0], a4_ii[1] = d1[0], d2[0], d1[1], d2[1]
    a4_jk[2] = a4_joint(d1, d2)
    if q3 is not None and np.real(q3[0]) > 1e-8:
        d3 = dist(q3, oracle)
        a2_i[2], a4_ii[2] = d3[0], d3[1]
        a4_jk[0] = a4_joint(d2, d3)
        a4_jk[1] = a4_joint(d1, d3)
    else:
        a2_i[2], a4_ii[2] = a2_orth(d1, d2), a4_orth(d1, d2)
        a4_jk[0] = a4_jointcross(d2, d1)
        a4_jk[1] = a4_jointcross(d1, d2)
    return a2_i, a4_ii, a4_jk


def _sym2(A):
    return (A + np.transpose(A, (2, 3, 0, 1))) / 2


def _sym4(A):
    # X(i,j,k,l) = (A(i,k,j,l) + A(k,j,i,l) + A(i,l,k,j) + A(l,j,k,i)) / 4
    return (np.einsum("ikjl->ijkl", A) + np.einsum("kjil->ijkl", A) + np.einsum("ilkj->ijkl", A)
            + np.einsum("ljki->ijkl", A)) / 4


def sachs_Q(a4_ii, a4_jk, Eg, n_grain=1):
    B = np.asarray(Eg, dtype=float) ** (2.0 / (n_grain + 1))
    lam = np.array([-B[0] + B[1] + B[2], B[0] - B[1] + B[2], B[0] + B[1] - B[2], B[3], B[4], B[5]])
    ci = np.concatenate([4.0 / 3 * lam[:3], 2 * lam[3:]])
    ji, ki = [1, 2, 0], [2, 0, 1]
    Q = np.zeros((3, 3, 3, 3))
    for i in range(3):
        M2 = (a4_ii[ji[i]] + a4_ii[ki[i]] - 2 * _sym2(a4_jk[i])) / 4
        Q += ci[i] * M2 + ci[i + 3] * _sym4(a4_jk[i])
    return Q


def _states(e1, e2, e3):
    e = [np.asarray(e1, float), np.asarray(e2, float), np.asarray(e3, float)]
    for (iv, iw) in [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]:
        v, w = e[iv], e[iw]
        tau = np.eye(3) / 3 - np.outer(v, v) if iv == iw else np.outer(v, w) + np.outer(w, v)
        yield tau, np.outer(v, w)


def Eij(q1, q2, q3, e1, e2, e3, Eg, oracle):
    iso = np.zeros(15, dtype=complex)
    iso[0] = 1 / np.sqrt(4 * np.pi)
    _, a4_ii, a4_jk = ai(q1, q2, q3, oracle)
    Q = sachs_Q(a4_ii, a4_jk, Eg)
    _, a4i, a4jki = ai(iso, iso, iso, oracle)
    Qi = sachs_Q(a4i, a4jki, Eg)
    d42 = lambda A, B: np.einsum("lkij,ji->lk", A, B)
    d22 = lambda A, B: np.einsum("ij,ji->", A, B)
    return np.array([d22(d42(Q, tau), vw) / d22(d42(Qi, tau), vw) for tau, vw in _states(e1, e2, e3)])


def _qnorm(a2_i, a4_ii, a4_jk, Lame):
    """ai_to_Qnorm_orthotropic (src/wavepropagation.f90:98-115), Voigt, k = z"""
    lam, mu = np.asarray(Lame[:6], float), np.asarray(Lame[6:], float)
    kk = np.zeros((3, 3))
    kk[2, 2] = 1
    Qn = np.zeros((3, 3))
    for i in range(3):
        a2 = a2_i[i]
        Qn += mu[i] / 2 * (a2 + a2 @ kk + kk @ a2 + np.einsum("ij,ji", a2, kk) * np.eye(3))
        Qn += lam[i] * np.einsum("lkij,ji->lk", a4_ii[i], kk)
        A = 2 * _sym2(a4_jk[i])
        Qn += lam[i + 3] * np.einsum("lkij,jl->ki", A, kk)
    return Qn


def Vi(q1, q2, q3, Lame, rho, theta, phi, oracle):
    """(ndir, 3) velocities (qS1, qS2, qP) per direction"""
    out = []
    for t, p in zip(np.atleast_1d(theta), np.atleast_1d(phi)):
        rot = lambda q: oracle.rotate_nlm4(np.asarray(q)[:15], -t, -p) if q is not None else None
        out.append(qnorm_to_vi(_qnorm(*ai(rot(q1), rot(q2), rot(q3), oracle), Lame), rho))
    return np.array(out)


# ---- discrete grains: (N, 3) axes b_i, n_i, v_i of N grains ----

def ai_discrete(bi, ni, vi):
    """ai_orthotropic_discrete (src/moments.f90:386-417): grain averages of
    q q, q q q q per axis and <q2 q2 q3 q3>, <q1 q1 q3 q3>, <q1 q1 q2 q2>"""
    q = [np.asarray(x, float) for x in (bi, ni, vi)]
    N = len(q[0])
    a2_i = np.array([np.einsum("gi,gj->ij", x, x) / N for x in q])
    a4_ii = np.array([np.einsum("gi,gj,gk,gl->ijkl", x, x, x, x) / N for x in q])
    jk = [(1, 2), (0, 2), (0, 1)]
    a4_jk = np.array([np.einsum("gi,gj,gk,gl->ijkl", q[j], q[j], q[k], q[k]) / N for j, k in jk])
    return a2_i, a4_ii, a4_jk


def Eij_discrete(bi, ni, vi, e1, e2, e3, Eg, oracle):
    """Eij_orthotropic_discrete (src/enhancementfactors.f90:191-266): the
    grains' moments in the Sachs tensor, normalised by the continuum
    isotropic fabric's"""
    iso = np.zeros(15, dtype=complex)
    iso[0] = 1 / np.sqrt(4 * np.pi)
    _, a4_ii, a4_jk = ai_discrete(bi, ni, vi)
    Q = sachs_Q(a4_ii, a4_jk, Eg)
    _, a4i, a4jki = ai(iso, iso, iso, oracle)
    Qi = sachs_Q(a4i, a4jki, Eg)
    d42 = lambda A, B: np.einsum("lkij,ji->lk", A, B)
    d22 = lambda A, B: np.einsum("ij,ji->", A, B)
    return np.array([d22(d42(Q, tau), vw) / d22(d42(Qi, tau), vw) for tau, vw in _states(e1, e2, e3)])


def rotate_vector(v, theta, phi):
    """src/rotation.f90:70-89: w = Ry(theta) Rz(phi) v"""
    c, s = np.cos(phi), np.sin(phi)
    Rz = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])
    c, s = np.cos(theta), np.sin(theta)
    Ry = np.array([[c, 0, s], [0, 1.0, 0], [-s, 0, c]])
    return np.asarray(v, float) @ (Ry @ Rz).T


def Vi_discrete(bi, ni, vi, Lame, rho, theta, phi):
    """Vi_elastic_orthotropic__discrete (src/wavepropagation.f90:70-96): the
    grains' axes rotated by (-theta, -phi) per direction, their moments, the
    Voigt acoustic tensor for k = z and the Cardano roots; (ndir, 3)"""
    out = []
    for t, p in zip(np.atleast_1d(theta), np.atleast_1d(phi)):
        r = [rotate_vector(x, -t, -p) for x in (bi, ni, vi)]
        out.append(qnorm_to_vi(_qnorm(*ai_discrete(*r), Lame), rho))
    return np.array(out)


def qnorm_to_vi(Q, rho):
    """(qS1, qS2, qP) of the acoustic tensor Q: Qnorm_to_vi with detQ_polycoefs,
    src/wavepropagation.f90:267-324 -- the Cardano roots of det(Q - rho v^2 I)
    in the slowness q^2 = 1/v^2 (omega = 1), principal complex powers with the
    REAL(4) exponents 2/3, 1/3, 4/3 the reference writes, v = 1/|sqrt(q^2)|."""
    w = rho
    I1 = np.trace(Q)
    I2 = (np.trace(Q) ** 2 - np.trace(Q @ Q)) / 2
    I3 = np.linalg.det(Q)
    r0, r2, r4, r6 = -w ** 3, I1 * w * w, -I2 * w, I3
    e23, e13, e43 = (float(np.float32(x)) for x in (2 / 3, 1 / 3, 4 / 3))
    a = complex((3 * r6 * r2 - r4 * r4) / (3 * r6 * r6))
    b = complex((2 * r4 ** 3 - 9 * r6 * r4 * r2 + 27 * r6 * r6 * r0) / (27 * r6 ** 3))
    root = np.sqrt(12 * a ** 3 + 81 * b ** 2 + 0j)
    denom = (6 + 0j) ** e23 * (-9 * b + root) ** e13
    D = (-9 * b + root) ** e23
    sh = complex(-r4 / (3 * r6))
    x1 = sh + (-2 * (3 + 0j) ** e13 * a + (2 + 0j) ** e13 * D) / denom
    x2 = sh + (-1 + 0j) ** e13 * (2 * (3 + 0j) ** e13 * a + (-2 + 0j) ** e13 * D) / denom
    x3 = sh + (-1 + 0j) ** e43 * (2 * (-3 + 0j) ** e13 * a + (2 + 0j) ** e13 * D) / denom
    return np.array([1 / abs(np.sqrt(x3)), 1 / abs(np.sqrt(x1)), 1 / abs(np.sqrt(x2))])
