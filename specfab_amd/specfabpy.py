"""Drop-in mirror of the reference's `specfabpy.specfab` module (the f2py
binding, src/specfabpy.f90) for the fabric-evolution and homogenization path,
backed by libspecfab_gpu on the GPU.

Usage mirrors the reference:

    from specfab_amd import specfabpy as sf
    lm, nlm_len = sf.init(12)
    M = sf.M_LROT(nlm, D, W, 1, 0) + sf.M_REG(nlm, D)

Like the reference, `init(L)` selects the module's active truncation; unlike
the reference, the library itself is reentrant (L travels with each call).
Inputs are numpy arrays in the reference's index order (3x3 tensors A[i, j],
nlm complex in specfab order); outputs are numpy arrays.  Batched, device-
resident work goes through `specfab_amd.batch.FabricBatch` instead.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, dptr, lib

_c = ctypes
_state = {"L": None, "n": None}

L2len = 6
L4len = 15
L6len = 28
L8len = 45
I20 = 3   # index of n_2^0 (0-based), src/specfabpy.f90:139-142
I40 = 10
I60 = 21
I80 = 36


def _ctx():
    return _lib.default_context().handle


def _L_of(nlm):
    n = len(nlm)
    L = int(round((-3 + np.sqrt(1 + 8 * n)) / 2))
    if (L + 1) * (L + 2) // 2 != n:
        raise ValueError("len(nlm)=%d is not an nlm length" % n)
    return L


def _m3(A):
    return np.ascontiguousarray(np.asarray(A, dtype=np.float64).reshape(3, 3))


def _z(nlm):
    return np.ascontiguousarray(np.asarray(nlm, dtype=np.complex128))


def _zp(a):
    return a.view(np.float64).ctypes.data_as(_lib._dp)


def init(L):
    """init(L) -> (lm (2, n), nlm_len): src/specfabpy.f90:150-160."""
    n = check_n(lib().sfg_nlm_len(int(L)))
    lm = np.zeros(2 * n, dtype=np.int32)
    check(lib().sfg_lm(int(L), lm.ctypes.data_as(_lib._ip)))
    _state["L"], _state["n"] = int(L), n
    _lib.default_context()
    return lm.reshape(n, 2).T.copy(), n


def check_n(n):
    if n < 0:
        check(n)
    return n


def nlm_len():
    return _state["n"]


def M_LROT(nlm, eps, omg, iota, zeta):
    """Lattice rotation operator, src/specfabpy.f90:172-180 -> dynamics.f90:52-97."""
    L = _L_of(nlm)
    n = len(nlm)
    M = np.zeros((n, n), dtype=np.complex128)
    check(lib().sfg_M_LROT(_ctx(), L, dptr(_m3(eps)), dptr(_m3(omg)), float(iota), float(zeta), _zp(M)))
    return M


def M_DDRX(nlm, tau):
    """DDRX operator (Gamma/Gamma0), src/specfabpy.f90:182-190 -> dynamics.f90:251-275."""
    L = _L_of(nlm)
    n = len(nlm)
    M = np.zeros((n, n), dtype=np.complex128)
    z = _z(nlm)
    check(lib().sfg_M_DDRX(_ctx(), L, _zp(z), dptr(_m3(tau)), _zp(M)))
    return M


def M_DDRX_src(nlm, tau):
    """DDRX source operator, src/specfabpy.f90:192-199 -> dynamics.f90:277-298."""
    L = _L_of(nlm)
    n = len(nlm)
    M = np.zeros((n, n), dtype=np.complex128)
    check(lib().sfg_M_DDRX_src(_ctx(), L, dptr(_m3(tau)), _zp(M)))
    return M


def M_CDRX(nlm):
    """CDRX operator (unscaled Laplacian), src/specfabpy.f90:228-235."""
    L = _L_of(nlm)
    n = len(nlm)
    M = np.zeros((n, n))
    check(lib().sfg_M_CDRX(_ctx(), L, dptr(M)))
    return M.astype(np.complex128)


def M_REG(nlm, eps):
    """Regularization operator, src/specfabpy.f90:237-245 -> dynamics.f90:494-518."""
    L = _L_of(nlm)
    n = len(nlm)
    M = np.zeros((n, n))
    check(lib().sfg_M_REG(_ctx(), L, dptr(_m3(eps)), dptr(M)))
    return M


def a2(nlm):
    """a2 = <c^2>, src/moments.f90:37-44."""
    L = _L_of(nlm)
    out = np.zeros(9)
    z = _z(nlm)
    check(lib().sfg_a2(_ctx(), L, _zp(z), dptr(out)))
    return out.reshape(3, 3)


def ev_D2(nlm, tau):
    L = _L_of(nlm)
    out = np.zeros(1)
    z = _z(nlm)
    check(lib().sfg_ev_D2(_ctx(), L, _zp(z), dptr(_m3(tau)), dptr(out)))
    return float(out[0])


def ugrad_to_D_and_W(ugrad):
    """src/deformationmodes.f90:29-35 (kinematics helper, host arithmetic only)."""
    G = np.asarray(ugrad, dtype=np.float64)
    return (G + G.T) / 2, (G - G.T) / 2


def eig(nlm):
    """eig(nlm) -> (ei (3,3) rows e1,e2,e3, lami (3)), largest first: src/frames.f90:14-22."""
    L = _L_of(nlm)
    ei = np.zeros(9)
    lam = np.zeros(3)
    z = _z(nlm)
    check(lib().sfg_eig(_ctx(), L, _zp(z), dptr(ei), dptr(lam)))
    return ei.reshape(3, 3), lam


def Eij_tranisotropic(nlm, e1, e2, e3, Eij_grain, alpha, n_grain):
    """Enhancement factors w.r.t. (e1,e2,e3): src/specfabpy.f90:390-400."""
    L = _L_of(nlm)
    out = np.zeros(6)
    z = _z(nlm)
    v = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    check(lib().sfg_Eij_tranisotropic(_ctx(), L, _zp(z), dptr(v(e1)), dptr(v(e2)), dptr(v(e3)), dptr(v(Eij_grain)),
                                      float(alpha), int(n_grain), dptr(out)))
    return out


def Eij_tranisotropic_arr(nlm, e1, e2, e3, Eij_grain, alpha, n_grain):
    """Batched Eij (src/specfabpy.f90:474-486): nlm (N, n), e_i (N, 3) -> (N, 6),
    evaluated in one GPU launch."""
    import torch
    from .batch import FabricBatch
    nlm = np.asarray(nlm)
    frames = np.stack([np.asarray(e1), np.asarray(e2), np.asarray(e3)], axis=1)
    b = FabricBatch(_L_of(nlm[0]), nlm=nlm)
    E = b.Eij(Eij_grain=Eij_grain, alpha=alpha, n_grain=n_grain, frames=frames)
    torch.cuda.synchronize()
    return E.cpu().numpy()


def a4(nlm):
    """a4 = <c^4> (3,3,3,3), src/moments.f90:46-55 (REAL(4) literals of the body)."""
    L = _L_of(nlm)
    out = np.zeros(81)
    z = _z(nlm)
    check(lib().sfg_a4(_ctx(), L, _zp(z), dptr(out)))
    return out.reshape(3, 3, 3, 3)


def Cij_to_Lame_tranisotropic(Cij):
    """(C11, C33, C55, C12, C13) -> (lam, mu, Elam, Emu, Egam), src/elasticities.f90:107-120."""
    C = np.ascontiguousarray(np.asarray(Cij, dtype=np.float64).reshape(5))
    out = np.zeros(5)
    check(lib().sfg_Cij_to_Lame_tranisotropic(dptr(C), dptr(out)))
    return out


def _dirs(theta, phi):
    th = np.ascontiguousarray(np.atleast_1d(np.asarray(theta, dtype=np.float64)).ravel())
    ph = np.ascontiguousarray(np.atleast_1d(np.asarray(phi, dtype=np.float64)).ravel())
    if th.size != ph.size:
        raise ValueError("theta and phi must have the same length")
    return th, ph


def Vi_elastic_tranisotropic(nlm, alpha, Lame_grain, rho, theta, phi):
    """Elastic phase velocities (qS1, qS2, qP) x ndir, src/specfabpy.f90:999-1008
    -> src/wavepropagation.f90:121-144.  Returns shape (3, ndir) like the reference."""
    L = _L_of(nlm)
    th, ph = _dirs(theta, phi)
    out = np.zeros((th.size, 3))
    Lm = np.ascontiguousarray(np.asarray(Lame_grain, dtype=np.float64).reshape(5))
    z = _z(nlm)
    check(lib().sfg_Vi_elastic_tranisotropic(_ctx(), L, _zp(z), float(alpha), dptr(Lm), float(rho), th.size,
                                             dptr(th), dptr(ph), dptr(out)))
    return out.T.copy()


def Vi_electromagnetic_tranisotropic(nlm, eps_m, eps_t, mu, theta, phi):
    """EM phase velocities (S1 fast, S2 slow) x ndir, src/specfabpy.f90:1024-1033
    -> src/wavepropagation.f90:217-261.  Returns shape (2, ndir)."""
    L = _L_of(nlm)
    th, ph = _dirs(theta, phi)
    out = np.zeros((th.size, 2))
    z = _z(nlm)
    check(lib().sfg_Vi_electromagnetic_tranisotropic(_ctx(), L, _zp(z), float(eps_m), float(eps_t), float(mu),
                                                     th.size, dptr(th), dptr(ph), dptr(out)))
    return out.T.copy()
