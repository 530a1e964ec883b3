"""ctypes face of oracle/_ref/libsfref.so (the reference's own FP64 Fortran,
built by oracle/ref.mk).  Test infrastructure: used by tests, the golden-vector
script and bench.py's CPU-baseline leg only, never by specfab_amd.

Array conventions mirror the Fortran: 3x3 tensors and n x n operators are
exchanged column-major; the helpers below take/return numpy arrays in natural
(row-major, C) index order and do the transposes.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "_ref", "libsfref.so")

_c = ctypes
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_cp = np.ctypeslib.ndpointer(dtype=np.complex128, flags="F_CONTIGUOUS")


def available():
    return os.path.exists(LIBPATH)


class RefLib:
    """One process = one active L (the reference keeps module globals)."""

    def __init__(self, path=LIBPATH):
        self.lib = _c.CDLL(path)
        L = self.lib
        L.sfref_init.argtypes = [_c.c_int]
        L.sfref_nlm_len.restype = _c.c_int
        L.sfref_M_LROT.argtypes = [_dp, _dp, _c.c_double, _c.c_double, _cp]
        L.sfref_M_DDRX.argtypes = [_cp, _dp, _cp]
        L.sfref_M_DDRX_src.argtypes = [_dp, _cp]
        L.sfref_M_REG.argtypes = [_dp, _dp]
        L.sfref_M_CDRX.argtypes = [_dp]
        L.sfref_rnlm_len.restype = _c.c_int
        L.sfref_reduce_M.argtypes = [_cp, _dp, _dp, _dp, _dp]
        L.sfref_nlm_to_rnlm.argtypes = [_cp, _cp]
        L.sfref_rnlm_to_nlm.argtypes = [_cp, _cp]
        L.sfref_ev_D2.argtypes = [_cp, _dp]
        L.sfref_ev_D2.restype = _c.c_double
        L.sfref_a2.argtypes = [_cp, _dp]
        L.sfref_a4.argtypes = [_cp, _dp]
        L.sfref_ev_ck_Mandel.argtypes = [_cp, _dp, _dp]
        L.sfref_eig.argtypes = [_cp, _dp, _dp]
        L.sfref_eig3.argtypes = [_dp, _dp, _dp]
        L.sfref_Evw_tranisotropic.argtypes = [_cp, _dp, _dp, _dp, _dp, _c.c_double, _c.c_int, _dp]
        L.sfref_Evw_orthotropic.argtypes = [_cp, _cp, _cp, _dp, _dp, _dp, _dp, _c.c_double, _c.c_int, _dp]
        L.sfref_eigframe.argtypes = [_dp, _c.c_int, _dp, _dp]
        L.sfref_nlm_LROT.argtypes = [_cp, _c.c_double, _c.c_int, _dp, _dp, _c.c_double, _cp]
        L.sfref_Eij_tranisotropic.argtypes = [_cp, _dp, _dp, _dp, _dp, _c.c_double, _c.c_int, _dp]
        L.sfref_Eij_eigframe_batch.argtypes = [_c.c_int, _cp, _dp, _c.c_double, _c.c_int, _dp, _dp]
        L.sfref_Cij_to_Lame_tranisotropic.argtypes = [_dp, _dp]
        L.sfref_Vi_elastic_tranisotropic.argtypes = [_cp, _c.c_double, _dp, _c.c_double, _c.c_int, _dp, _dp, _dp]
        L.sfref_Qnorm_tranisotropic.argtypes = [_cp, _c.c_double, _dp, _dp]
        L.sfref_Vi_electromagnetic_tranisotropic.argtypes = [_cp, _c.c_double, _c.c_double, _c.c_double,
                                                             _c.c_int, _dp, _dp, _dp]
        L.sfref_rotate_nlm4.argtypes = [_cp, _c.c_double, _c.c_double, _cp]
        L.sfref_rotate_nlm.argtypes = [_cp, _c.c_double, _c.c_double, _cp]
        L.sfref_a4_eigentensors.argtypes = [_cp, _dp, _dp]
        L.sfref_nlm_isvalid.argtypes = [_c.c_int, _cp, _cp, np.ctypeslib.ndpointer(dtype=np.int32)]
        L.sfref_ortho_moments.argtypes = [_cp, _cp, _dp, _dp, _dp, _dp]
        L.sfref_Eij_orthotropic.argtypes = [_cp, _cp, _cp, _dp, _dp, _dp, _dp, _c.c_double, _c.c_int, _dp]
        L.sfref_Eij_orthotropic_discrete.argtypes = [_c.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _c.c_double,
                                                     _c.c_int, _dp]
        L.sfref_Vi_elastic_orthotropic.argtypes = [_cp, _cp, _cp, _c.c_double, _dp, _c.c_double, _c.c_int, _dp, _dp,
                                                   _dp]
        L.sfref_Vi_elastic_orthotropic_discrete.argtypes = [_c.c_int, _dp, _c.c_double, _dp, _c.c_double, _c.c_int,
                                                            _dp, _dp, _dp]
        L.sfref_ev_ck.argtypes = [_cp, _dp, _dp, _dp, _dp]
        L.sfref_ev_D.argtypes = [_cp, _dp, _c.c_int, _dp]
        L.sfref_E_CAFFE.argtypes = [_cp, _dp, _c.c_double, _c.c_double, _c.c_int, _dp]
        L.sfref_ri_LROT.argtypes = [_c.c_int, _c.c_int, _dp, _c.c_double, _dp, _dp, _c.c_double, _dp]
        L.sfref_ri_to_nlm.argtypes = [_c.c_int, _c.c_int, _dp, _dp, _cp]
        L.sfref_evolve.argtypes = [_c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_double, _c.c_double,
                                   _c.c_double, _dp, _dp, _dp, _cp]
        L.sfref_a2_to_nlm.argtypes = [_dp, _cp]
        L.sfref_a4_to_nlm.argtypes = [_dp, _cp]
        L.sfref_Gamma0.argtypes = [_dp, _c.c_double, _c.c_double, _c.c_double]
        L.sfref_Gamma0.restype = _c.c_double
        L.sfref_apply_bounds.argtypes = [_cp, _cp]
        L.sfref_pureshear_F.argtypes = [_c.c_int, _c.c_double, _c.c_double, _c.c_double, _dp]
        L.sfref_pureshear_ugrad.argtypes = [_c.c_int, _c.c_double, _c.c_double, _dp]
        L.sfref_pureshear_strainii_to_t.argtypes = [_c.c_double, _c.c_double]
        L.sfref_pureshear_strainii_to_t.restype = _c.c_double
        L.sfref_simpleshear_F.argtypes = [_c.c_int, _c.c_double, _c.c_double, _dp]
        L.sfref_simpleshear_ugrad.argtypes = [_c.c_int, _c.c_double, _dp]
        L.sfref_simpleshear_gamma_to_t.argtypes = [_c.c_double, _c.c_double]
        L.sfref_simpleshear_gamma_to_t.restype = _c.c_double
        self.L = None
        self.n = 0

    # -- helpers --
    @staticmethod
    def _m3(A):
        return np.asfortranarray(np.asarray(A, dtype=np.float64).reshape(3, 3))

    @staticmethod
    def _v(x, dtype=np.float64):
        return np.asfortranarray(np.asarray(x, dtype=dtype).ravel())

    def init(self, L):
        self.lib.sfref_init(int(L))
        self.L = int(L)
        self.n = self.lib.sfref_nlm_len()
        return self.n

    def M_LROT(self, eps, omg, iota, zeta):
        M = np.zeros((self.n, self.n), dtype=np.complex128, order="F")
        self.lib.sfref_M_LROT(self._m3(eps), self._m3(omg), float(iota), float(zeta), M)
        return np.ascontiguousarray(M)

    def M_DDRX(self, nlm, tau):
        M = np.zeros((self.n, self.n), dtype=np.complex128, order="F")
        self.lib.sfref_M_DDRX(self._v(nlm, np.complex128), self._m3(tau), M)
        return np.ascontiguousarray(M)

    def M_DDRX_src(self, tau):
        M = np.zeros((self.n, self.n), dtype=np.complex128, order="F")
        self.lib.sfref_M_DDRX_src(self._m3(tau), M)
        return np.ascontiguousarray(M)

    def M_REG(self, D):
        M = np.zeros((self.n, self.n), order="F")
        self.lib.sfref_M_REG(self._m3(D), M)
        return np.ascontiguousarray(M)

    def M_CDRX(self):
        M = np.zeros((self.n, self.n), order="F")
        self.lib.sfref_M_CDRX(M)
        return np.ascontiguousarray(M)

    # reduced form (src/reducedform.f90:76-165); Mrr.. in natural [row, col] order
    def reduce_M(self, M):
        r = self.lib.sfref_rnlm_len()
        out = [np.zeros((r, r), order="F") for _ in range(4)]
        self.lib.sfref_reduce_M(np.asfortranarray(np.asarray(M, dtype=np.complex128)), *out)
        return tuple(np.ascontiguousarray(o) for o in out)

    def nlm_to_rnlm(self, nlm):
        r = np.zeros(self.lib.sfref_rnlm_len(), dtype=np.complex128)
        self.lib.sfref_nlm_to_rnlm(self._v(nlm, np.complex128), r)
        return r

    def rnlm_to_nlm(self, rnlm):
        nlm = np.zeros(self.n, dtype=np.complex128)
        self.lib.sfref_rnlm_to_nlm(self._v(rnlm, np.complex128), nlm)
        return nlm

    def ev_D2(self, nlm, tau):
        return self.lib.sfref_ev_D2(self._v(nlm, np.complex128), self._m3(tau))

    def a2(self, nlm):
        out = np.zeros((3, 3), order="F")
        self.lib.sfref_a2(self._v(nlm, np.complex128), out)
        return np.ascontiguousarray(out)

    def a4(self, nlm):
        out = np.zeros((3, 3, 3, 3), order="F")
        self.lib.sfref_a4(self._v(nlm, np.complex128), out)
        return np.ascontiguousarray(out)

    def ev_ck_Mandel(self, nlm):
        a2v = np.zeros(6, order="F")
        a4v = np.zeros((6, 6), order="F")
        self.lib.sfref_ev_ck_Mandel(self._v(nlm, np.complex128), a2v, a4v)
        return a2v, np.ascontiguousarray(a4v)

    def eig(self, nlm):
        ei = np.zeros((3, 3), order="F")
        lami = np.zeros(3, order="F")
        self.lib.sfref_eig(self._v(nlm, np.complex128), ei, lami)
        return np.ascontiguousarray(ei), lami.copy()

    def Evw_tranisotropic(self, nlm, v, w, tau, Eg, alpha, ng):
        out = np.zeros(1, order="F")
        self.lib.sfref_Evw_tranisotropic(self._v(nlm, np.complex128), self._v(v), self._v(w),
                                         np.asfortranarray(tau, dtype=np.float64), self._v(Eg), float(alpha), int(ng),
                                         out)
        return float(out[0])

    def Evw_orthotropic(self, q1, q2, q3, v, w, tau, Eg, alpha, ng):
        out = np.zeros(1, order="F")
        self.lib.sfref_Evw_orthotropic(self._v(q1, np.complex128), self._v(q2, np.complex128),
                                       self._v(q3, np.complex128), self._v(v), self._v(w),
                                       np.asfortranarray(tau, dtype=np.float64), self._v(Eg), float(alpha), int(ng), out)
        return float(out[0])

    def eig3(self, M):
        """-> ei (3, 3) rows e1, e2, e3, lami (3)"""
        ei = np.zeros((3, 3), order="F")
        lami = np.zeros(3, order="F")
        self.lib.sfref_eig3(np.asfortranarray(M, dtype=np.float64), ei, lami)
        return np.ascontiguousarray(ei), lami.copy()

    def eigframe(self, M, plane):
        ei = np.zeros((3, 3), order="F")
        lami = np.zeros(3, order="F")
        self.lib.sfref_eigframe(np.asfortranarray(M, dtype=np.float64), {"ij": 0, "xy": 1, "xz": 2}[plane], ei, lami)
        return np.ascontiguousarray(ei), lami.copy()

    def nlm_LROT(self, nlm0, dt, Nt, D, W, iota):
        out = np.zeros((Nt, len(nlm0)), dtype=np.complex128, order="F")
        self.lib.sfref_nlm_LROT(self._v(nlm0, np.complex128), float(dt), int(Nt),
                                np.asfortranarray(D, dtype=np.float64), np.asfortranarray(W, dtype=np.float64),
                                float(iota), out)
        return np.ascontiguousarray(out)

    def Eij_tranisotropic(self, nlm, e1, e2, e3, Eij_grain, alpha, n_grain):
        out = np.zeros(6, order="F")
        self.lib.sfref_Eij_tranisotropic(self._v(nlm, np.complex128), self._v(e1), self._v(e2), self._v(e3),
                                         self._v(Eij_grain), float(alpha), int(n_grain), out)
        return out.copy()

    def Eij_eigframe_batch(self, nlm, Eij_grain, alpha, n_grain):
        """nlm (P, n) -> Eij (P, 6), lami (P, 3)."""
        P = nlm.shape[0]
        X = np.asfortranarray(np.asarray(nlm, dtype=np.complex128).T)
        E = np.zeros((6, P), order="F")
        lam = np.zeros((3, P), order="F")
        self.lib.sfref_Eij_eigframe_batch(P, X, self._v(Eij_grain), float(alpha), int(n_grain), E, lam)
        return np.ascontiguousarray(E.T), np.ascontiguousarray(lam.T)

    def Cij_to_Lame_tranisotropic(self, Cij):
        out = np.zeros(5, order="F")
        self.lib.sfref_Cij_to_Lame_tranisotropic(self._v(Cij), out)
        return out.copy()

    def Vi_elastic_tranisotropic(self, nlm, alpha, Lame, rho, theta, phi):
        theta = self._v(np.atleast_1d(theta))
        phi = self._v(np.atleast_1d(phi))
        out = np.zeros((3, theta.size), order="F")
        self.lib.sfref_Vi_elastic_tranisotropic(self._v(nlm, np.complex128), float(alpha), self._v(Lame),
                                                float(rho), theta.size, theta, phi, out)
        return np.ascontiguousarray(out)

    def rotate_nlm4(self, nlm4, theta, phi):
        out = np.zeros(15, dtype=np.complex128, order="F")
        self.lib.sfref_rotate_nlm4(self._v(nlm4[:15], np.complex128), float(theta), float(phi), out)
        return out.copy()

    def rotate_nlm(self, nlm, theta, phi):
        out = np.zeros(self.n, dtype=np.complex128, order="F")
        self.lib.sfref_rotate_nlm(self._v(nlm, np.complex128), float(theta), float(phi), out)
        return out.copy()

    def a4_eigentensors(self, nlm):
        """(Q (6, 3, 3): eigentensors of the Mandel a4, eigvals (6) ascending)"""
        Q = np.zeros((3, 3, 6), order="F")
        ev = np.zeros(6, order="F")
        self.lib.sfref_a4_eigentensors(self._v(nlm, np.complex128), Q, ev)
        return np.ascontiguousarray(np.transpose(Q, (2, 0, 1))), ev.copy()

    def nlm_isvalid(self, nhat20, nhat40):
        n20 = np.ascontiguousarray(nhat20, dtype=np.complex128)
        n40 = np.ascontiguousarray(nhat40, dtype=np.complex128)
        ok = np.zeros(len(n20), dtype=np.int32)
        self.lib.sfref_nlm_isvalid(len(n20), n20, n40, ok)
        return ok.astype(bool)

    def ortho_moments(self, b, n):
        """(a4_joint, a4_jointcross, a2_orth, a4_orth) of the distributions b, n"""
        out = [np.zeros((3,) * k, order="F") for k in (4, 4, 2, 4)]
        self.lib.sfref_ortho_moments(self._v(b, np.complex128), self._v(n, np.complex128), *out)
        return tuple(np.ascontiguousarray(o) for o in out)

    def Eij_orthotropic(self, q1, q2, q3, e1, e2, e3, Eg, alpha, ng):
        out = np.zeros(6, order="F")
        self.lib.sfref_Eij_orthotropic(self._v(q1, np.complex128), self._v(q2, np.complex128),
                                       self._v(q3, np.complex128), self._v(e1), self._v(e2), self._v(e3),
                                       self._v(Eg), float(alpha), int(ng), out)
        return out.copy()

    def Eij_orthotropic_discrete(self, bi, ni, vi, e1, e2, e3, Eg, alpha, ng):
        out = np.zeros(6, order="F")
        f = lambda a: np.asfortranarray(a, dtype=np.float64)
        self.lib.sfref_Eij_orthotropic_discrete(len(bi), f(bi), f(ni), f(vi), self._v(e1), self._v(e2), self._v(e3),
                                                self._v(Eg), float(alpha), int(ng), out)
        return out.copy()

    def Vi_elastic_orthotropic(self, q1, q2, q3, alpha, Lame, rho, theta, phi):
        theta = self._v(np.atleast_1d(theta))
        phi = self._v(np.atleast_1d(phi))
        out = np.zeros((3, theta.size), order="F")
        self.lib.sfref_Vi_elastic_orthotropic(self._v(q1, np.complex128), self._v(q2, np.complex128),
                                              self._v(q3, np.complex128), float(alpha), self._v(Lame), float(rho),
                                              theta.size, theta, phi, out)
        return np.ascontiguousarray(out)

    def Vi_elastic_orthotropic_discrete(self, mi, alpha, Lame, rho, theta, phi):
        """mi (3, 3, N) = (m'_i, xyz, grain), as the reference takes it"""
        theta = self._v(np.atleast_1d(theta))
        phi = self._v(np.atleast_1d(phi))
        out = np.zeros((3, theta.size), order="F")
        m = np.asfortranarray(mi, dtype=np.float64)
        self.lib.sfref_Vi_elastic_orthotropic_discrete(m.shape[2], m, float(alpha), self._v(Lame), float(rho),
                                                       theta.size, theta, phi, out)
        return np.ascontiguousarray(out)

    def ev_ck(self, nlm):
        """(a2, a4, a6, a8) full tensors of f_ev_ck(nlm, 'f') (L >= 8)"""
        out = [np.zeros((3,) * k, order="F") for k in (2, 4, 6, 8)]
        self.lib.sfref_ev_ck(self._v(nlm, np.complex128), *out)
        return tuple(np.ascontiguousarray(o) for o in out)

    def ev_D(self, nlm, tau, n_rss):
        out = np.zeros(1, order="F")
        self.lib.sfref_ev_D(self._v(nlm, np.complex128), self._m3(tau), int(n_rss), out)
        return float(out[0])

    def E_CAFFE(self, nlm, eps, Emin, Emax, n_grain):
        out = np.zeros(1, order="F")
        self.lib.sfref_E_CAFFE(self._v(nlm, np.complex128), self._m3(eps), float(Emin), float(Emax), int(n_grain),
                               out)
        return float(out[0])

    def ri_LROT(self, ri0, dt, Nt, D, W, iota):
        """ri0 (N, 3), D and W (Nt, 3, 3) -> ri (Nt, N, 3)"""
        r0 = np.asfortranarray(ri0, dtype=np.float64)
        N = r0.shape[0]
        out = np.zeros((Nt, N, 3), order="F")
        self.lib.sfref_ri_LROT(N, int(Nt), r0, float(dt), np.asfortranarray(D, dtype=np.float64),
                               np.asfortranarray(W, dtype=np.float64), float(iota), out)
        return np.ascontiguousarray(out)

    def ri_to_nlm(self, ri, wi, L):
        """ri (N, 3), wi (N) -> nlm ((L+1)(L+2)/2) complex"""
        r = np.asfortranarray(ri, dtype=np.float64)
        out = np.zeros((L + 1) * (L + 2) // 2, dtype=np.complex128, order="F")
        self.lib.sfref_ri_to_nlm(r.shape[0], int(L), r, self._v(wi), out)
        return out.copy()

    def Qnorm_tranisotropic(self, nlm, alpha, Lame):
        out = np.zeros((3, 3), order="F")
        self.lib.sfref_Qnorm_tranisotropic(self._v(nlm, np.complex128), float(alpha), self._v(Lame), out)
        return np.ascontiguousarray(out)

    def Vi_electromagnetic_tranisotropic(self, nlm, eps_m, eps_t, mu, theta, phi):
        theta = self._v(np.atleast_1d(theta))
        phi = self._v(np.atleast_1d(phi))
        out = np.zeros((2, theta.size), order="F")
        self.lib.sfref_Vi_electromagnetic_tranisotropic(self._v(nlm, np.complex128), float(eps_m), float(eps_t),
                                                        float(mu), theta.size, theta, phi, out)
        return np.ascontiguousarray(out)

    def evolve(self, nlm, ugrad, gam0, lam, nsteps, dt, scheme="rk4", procs=("LROT", "REG"),
               iota=1.0, zeta=0.0):
        """nlm (P, n) complex, ugrad (P, 3, 3) -> advanced nlm (P, n)."""
        bits = {"LROT": 1, "DDRX": 2, "CDRX": 4, "REG": 8}
        mask = 0
        for p in procs:
            mask |= bits[p]
        P = nlm.shape[0]
        X = np.array(np.asarray(nlm, dtype=np.complex128).T, order="F", copy=True)   # advanced in place
        U = np.asfortranarray(np.transpose(np.asarray(ugrad, dtype=np.float64), (1, 2, 0)))
        g = self._v(np.broadcast_to(gam0, (P,)))
        lm = self._v(np.broadcast_to(lam, (P,)))
        self.lib.sfref_evolve(P, int(nsteps), 0 if scheme == "euler" else 1, mask, float(dt), float(iota),
                              float(zeta), U, g, lm, X)
        return np.ascontiguousarray(X.T)

    # -- initial states, rate factor, bounds (src/moments.f90:68-84,
    #    src/dynamics.f90:451-472, 530-557) --
    def a2_to_nlm(self, a2):
        out = np.zeros(6, dtype=np.complex128, order="F")
        self.lib.sfref_a2_to_nlm(self._m3(a2), out)
        return out.copy()

    def a4_to_nlm(self, a4):
        out = np.zeros(15, dtype=np.complex128, order="F")
        self.lib.sfref_a4_to_nlm(np.asfortranarray(np.asarray(a4, dtype=np.float64).reshape(3, 3, 3, 3)), out)
        return out.copy()

    def Gamma0(self, eps, T, A, Q):
        return self.lib.sfref_Gamma0(self._m3(eps), float(T), float(A), float(Q))

    def apply_bounds(self, nlm):
        out = np.zeros(self.n, dtype=np.complex128, order="F")
        self.lib.sfref_apply_bounds(self._v(nlm, np.complex128), out)
        return out.copy()

    # -- kinematics (src/deformationmodes.f90) --
    def pureshear_F(self, ax, q, tau, t):
        out = np.zeros((3, 3), order="F")
        self.lib.sfref_pureshear_F(int(ax), float(q), float(tau), float(t), out)
        return np.ascontiguousarray(out)

    def pureshear_ugrad(self, ax, q, tau):
        out = np.zeros((3, 3), order="F")
        self.lib.sfref_pureshear_ugrad(int(ax), float(q), float(tau), out)
        return np.ascontiguousarray(out)

    def pureshear_strainii_to_t(self, s, tau):
        return self.lib.sfref_pureshear_strainii_to_t(float(s), float(tau))

    def simpleshear_F(self, plane, tau, t):
        out = np.zeros((3, 3), order="F")
        self.lib.sfref_simpleshear_F(int(plane), float(tau), float(t), out)
        return np.ascontiguousarray(out)

    def simpleshear_ugrad(self, plane, tau):
        out = np.zeros((3, 3), order="F")
        self.lib.sfref_simpleshear_ugrad(int(plane), float(tau), out)
        return np.ascontiguousarray(out)

    def simpleshear_gamma_to_t(self, gam, tau):
        return self.lib.sfref_simpleshear_gamma_to_t(float(gam), float(tau))
