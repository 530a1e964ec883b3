"""ctypes face of the C oracle oracle/_build/libsforacle.so (oracle/sf_oracle.c).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  Arrays use natural numpy
(row-major) order; nlm is complex128 in specfab coefficient order.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "_build", "libsforacle.so")

_c = ctypes
_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_z = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")

PROC_BITS = {"LROT": 1, "DDRX": 2, "CDRX": 4, "REG": 8}


def nlm_len(L):
    return (L + 1) * (L + 2) // 2


def lm_list(L):
    return [(l, m) for l in range(0, L + 1, 2) for m in range(-l, l + 1)]


class Oracle:
    def __init__(self, path=LIBPATH):
        if not os.path.exists(path):
            raise RuntimeError("C oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        self.lib = lib = _c.CDLL(path)
        lib.sfo_nlm_len.argtypes = [_c.c_int]
        lib.sfo_nlm_len.restype = _c.c_int
        lib.sfo_gaunt_count.restype = _c.c_int
        lib.sfo_M_LROT.argtypes = [_c.c_int, _d, _d, _c.c_double, _c.c_double, _z]
        lib.sfo_M_DDRX_src.argtypes = [_c.c_int, _d, _z]
        lib.sfo_M_DDRX.argtypes = [_c.c_int, _z, _d, _z]
        lib.sfo_ddrx_weights.argtypes = [_d, _z]
        lib.sfo_ev_D2.argtypes = [_z, _d]
        lib.sfo_ev_D2.restype = _c.c_double
        lib.sfo_a2.argtypes = [_z, _d]
        lib.sfo_a4_mandel.argtypes = [_z, _d]
        lib.sfo_ev_ck_Mandel.argtypes = [_z, _d, _d]
        lib.sfo_M_REG_diag.argtypes = [_c.c_int, _d, _d]
        lib.sfo_M_REG_diag.restype = _c.c_int
        lib.sfo_M_CDRX_diag.argtypes = [_c.c_int, _d]
        lib.sfo_evolve.argtypes = [_c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_double, _c.c_double,
                                   _c.c_double, _d, _d, _d, _z]
        lib.sfo_evolve.restype = _c.c_int
        self._extra()

    def _extra(self):
        """Homogenization entry points (registered where the oracle has them)."""
        lib = self.lib
        sig = {
            "sfo_eig": [_z, _d, _d],
            "sfo_eig3": [_d, _d, _d],
            "sfo_Eij_tranisotropic": [_z, _d, _d, _d, _d, _c.c_double, _c.c_int, _d],
            "sfo_a4": [_z, _d],
            "sfo_Cij_to_Lame_tranisotropic": [_d, _d],
            "sfo_Vi_elastic_tranisotropic": [_z, _c.c_double, _d, _c.c_double, _c.c_int, _d, _d, _d],
            "sfo_Vi_electromagnetic_tranisotropic": [_z, _c.c_double, _c.c_double, _c.c_double, _c.c_int, _d, _d, _d],
            "sfo_rotate_nlm4": [_z, _c.c_double, _c.c_double, _z],
            "sfo_Qnorm_tranisotropic": [_z, _c.c_double, _d, _d],
        }
        for name, args in sig.items():
            if hasattr(lib, name):
                getattr(lib, name).argtypes = args
                getattr(lib, name).restype = _c.c_int if name in ("sfo_Eij_tranisotropic", "sfo_Vi_elastic_tranisotropic",
                                                                 "sfo_Qnorm_tranisotropic") else None

    @staticmethod
    def _m3(A):
        return np.ascontiguousarray(np.asarray(A, dtype=np.float64).reshape(3, 3))

    @staticmethod
    def _z(x):
        return np.ascontiguousarray(np.asarray(x, dtype=np.complex128))

    def M_LROT(self, L, eps, omg, iota, zeta):
        n = nlm_len(L)
        M = np.zeros((n, n), dtype=np.complex128)
        self.lib.sfo_M_LROT(L, self._m3(eps), self._m3(omg), float(iota), float(zeta), M)
        return M

    def M_DDRX_src(self, L, tau):
        n = nlm_len(L)
        M = np.zeros((n, n), dtype=np.complex128)
        self.lib.sfo_M_DDRX_src(L, self._m3(tau), M)
        return M

    def M_DDRX(self, L, nlm, tau):
        n = nlm_len(L)
        M = np.zeros((n, n), dtype=np.complex128)
        self.lib.sfo_M_DDRX(L, self._z(nlm), self._m3(tau), M)
        return M

    def ddrx_weights(self, tau):
        g = np.zeros(15, dtype=np.complex128)
        self.lib.sfo_ddrx_weights(self._m3(tau), g)
        return g

    def ev_D2(self, nlm, tau):
        return self.lib.sfo_ev_D2(self._z(nlm), self._m3(tau))

    def a2(self, nlm):
        out = np.zeros(9)
        self.lib.sfo_a2(self._z(nlm), out)
        return out.reshape(3, 3)

    def ev_ck_Mandel(self, nlm):
        a2v = np.zeros(6)
        a4v = np.zeros(36)
        self.lib.sfo_ev_ck_Mandel(self._z(nlm), a2v, a4v)
        return a2v, a4v.reshape(6, 6)

    def M_REG(self, L, D):
        d = np.zeros(nlm_len(L))
        if self.lib.sfo_M_REG_diag(L, self._m3(D), d) != 0:
            raise ValueError("M_REG undefined for L=%d" % L)
        return np.diag(d)

    def M_CDRX(self, L):
        d = np.zeros(nlm_len(L))
        self.lib.sfo_M_CDRX_diag(L, d)
        return np.diag(d)

    def evolve(self, L, nlm, ugrad, gam0, lam, nsteps, dt, scheme="rk4", procs=("LROT", "REG"),
               iota=1.0, zeta=0.0):
        nlm = np.array(nlm, dtype=np.complex128, order="C")
        P = nlm.shape[0]
        mask = 0
        for p in procs:
            mask |= PROC_BITS[p]
        U = np.ascontiguousarray(np.asarray(ugrad, dtype=np.float64).reshape(P, 3, 3))
        g = np.ascontiguousarray(np.broadcast_to(np.asarray(gam0, dtype=np.float64), (P,)))
        lm = np.ascontiguousarray(np.broadcast_to(np.asarray(lam, dtype=np.float64), (P,)))
        rc = self.lib.sfo_evolve(L, P, int(nsteps), 0 if scheme == "euler" else 1, mask, float(dt), float(iota),
                                 float(zeta), U, g, lm, nlm)
        if rc != 0:
            raise ValueError("sfo_evolve failed (rc=%d)" % rc)
        return nlm

    # -- homogenization --
    def eig(self, nlm):
        ei = np.zeros(9)
        lam = np.zeros(3)
        self.lib.sfo_eig(self._z(nlm), ei, lam)
        return ei.reshape(3, 3), lam

    def Eij_tranisotropic(self, nlm, e1, e2, e3, Eij_grain, alpha, n_grain):
        out = np.zeros(6)
        v = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64))
        self.lib.sfo_Eij_tranisotropic(self._z(nlm), v(e1), v(e2), v(e3), v(Eij_grain), float(alpha),
                                       int(n_grain), out)
        return out

    def a4(self, nlm):
        out = np.zeros(81)
        self.lib.sfo_a4(self._z(nlm), out)
        return out.reshape(3, 3, 3, 3)

    def Cij_to_Lame_tranisotropic(self, Cij):
        out = np.zeros(5)
        self.lib.sfo_Cij_to_Lame_tranisotropic(np.ascontiguousarray(Cij, dtype=np.float64), out)
        return out

    def Vi_elastic_tranisotropic(self, nlm, alpha, Lame, rho, theta, phi):
        th = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
        ph = np.ascontiguousarray(np.atleast_1d(phi), dtype=np.float64)
        out = np.zeros((th.size, 3))
        self.lib.sfo_Vi_elastic_tranisotropic(self._z(nlm), float(alpha), np.ascontiguousarray(Lame, dtype=np.float64),
                                              float(rho), th.size, th, ph, out)
        return out.T.copy()

    def Qnorm_tranisotropic(self, nlm, alpha, Lame):
        out = np.zeros(9)
        self.lib.sfo_Qnorm_tranisotropic(self._z(nlm), float(alpha), np.ascontiguousarray(Lame, dtype=np.float64), out)
        return out.reshape(3, 3)

    def rotate_nlm4(self, nlm4, theta, phi):
        out = np.zeros(15, dtype=np.complex128)
        self.lib.sfo_rotate_nlm4(self._z(nlm4[:15]), float(theta), float(phi), out)
        return out

    def Vi_electromagnetic_tranisotropic(self, nlm, eps_m, eps_t, mu, theta, phi):
        th = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
        ph = np.ascontiguousarray(np.atleast_1d(phi), dtype=np.float64)
        out = np.zeros((th.size, 2))
        self.lib.sfo_Vi_electromagnetic_tranisotropic(self._z(nlm), float(eps_m), float(eps_t), float(mu),
                                                      th.size, th, ph, out)
        return out.T.copy()


# ---- reduced real form (numpy; index and sign work only) ----------------------
# Restates src/reducedform.f90: rnlm holds the m >= 0 coefficients of every even
# l in the order (n_0^0, n_2^0, n_2^1, n_2^2, n_4^0, ...) (:37-44, rlm at :52),
# and a real ODF has n_l^{-m} = (-1)^m conj(n_l^m).

def rnlm_len(L):
    """(L+2)^2/4 (src/reducedform.f90:25)."""
    return (L + 2) ** 2 // 4


def rlm_list(L):
    return [(l, m) for l in range(0, L + 1, 2) for m in range(0, l + 1)]


def _full_pos(L):
    """0-based nlm position of (l, m) (the I_all index list, src/reducedform.f90:37)."""
    pos = {lm: k for k, lm in enumerate(lm_list(L))}
    return pos


def nlm_to_rnlm(nlm, L):
    """nlm_to_rnlm (src/reducedform.f90:176-183): rnlm = nlm at the m >= 0 positions."""
    pos = _full_pos(L)
    return np.asarray(nlm, dtype=np.complex128)[..., [pos[lm] for lm in rlm_list(L)]]


def rnlm_to_nlm(rnlm, L):
    """rnlm_to_nlm (src/reducedform.f90:160-174): m >= 0 copied, m < 0 from
    (-1)^m conj(n_l^{|m|}) (the Mdiag factor and the conjugation at I_neg)."""
    rnlm = np.asarray(rnlm, dtype=np.complex128)
    ridx = {lm: k for k, lm in enumerate(rlm_list(L))}
    out = np.zeros(rnlm.shape[:-1] + (nlm_len(L),), dtype=np.complex128)
    for j, (l, m) in enumerate(lm_list(L)):
        v = rnlm[..., ridx[(l, abs(m))]]
        out[..., j] = v if m >= 0 else (-1) ** m * np.conj(v)
    return out


def reduce_M(M, L):
    """reduce_M (src/reducedform.f90:76-120): for reduced column (l, m) the +m
    and -m columns of M at the reduced rows, Zp and Zn, combine with
    s = (-1)^m (s = 0 for m = 0) into Mrr = Re Zp + s Re Zn,
    Mri = -Im Zp + s Im Zn, Mir = Im Zp + s Im Zn, Mii = Re Zp - s Re Zn."""
    M = np.asarray(M, dtype=np.complex128)
    pos = _full_pos(L)
    rows = [pos[lm] for lm in rlm_list(L)]
    r = len(rows)
    out = [np.zeros(M.shape[:-2] + (r, r)) for _ in range(4)]
    for jj, (l, m) in enumerate(rlm_list(L)):
        zp = M[..., rows, pos[(l, m)]]
        zn = M[..., rows, pos[(l, -m)]]
        s = 0 if m == 0 else (-1) ** m
        out[0][..., jj] = zp.real + s * zn.real
        out[1][..., jj] = -zp.imag + s * zn.imag
        out[2][..., jj] = zp.imag + s * zn.imag
        out[3][..., jj] = zp.real - s * zn.real
    return tuple(out)
