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


def nlm_LROT(nlm0, dt, Nt, D, W, iota):
    """Euler trajectory (Nt, n) of the lattice rotation with per-step strain
    rate D and spin W (Nt, 3, 3): src/specfabpy.f90:260-269 ->
    src/dynamics.f90:99-110 (nlm[0] = nlm0, then Nt - 1 steps; zeta = 0).
    Each step is one launch of the stepping kernel with velocity gradient
    D_j + W_j (D_j symmetric, W_j antisymmetric, as the reference's
    kinematics give them)."""
    import torch
    from .batch import FabricBatch
    z0 = _z(nlm0)
    L = _L_of(z0)
    Nt = int(Nt)
    D = np.asarray(D, dtype=np.float64)
    W = np.asarray(W, dtype=np.float64)
    if Nt < 1 or D.shape[0] < Nt - 1 or W.shape[0] < Nt - 1 or D.shape[1:] != (3, 3) or W.shape[1:] != (3, 3):
        raise ValueError("nlm_LROT: D and W must be (>= Nt-1, 3, 3)")
    fb = FabricBatch(L, nlm=z0[None], ugrad=np.zeros((1, 3, 3)))
    out = np.zeros((Nt, len(z0)), dtype=np.complex128)
    out[0] = z0
    G = torch.as_tensor(D + W, device=fb.dev)
    for j in range(Nt - 1):
        fb.ugrad.copy_(G[j].reshape(1, 3, 3))
        fb.evolve(1, dt, scheme="euler", procs=("LROT",), iota=iota, zeta=0.0)
        out[j + 1] = fb.nlm()[0].cpu().numpy()
    return out


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


def _L_of_rnlm_len(rn):
    h = int(round(np.sqrt(rn)))
    if h < 1 or h * h != rn:
        raise ValueError("rnlm_len=%d is not a reduced-form length" % rn)
    return 2 * h - 2


def get_rnlm_len():
    """rnlm_len of the L set by init(), src/specfabpy.f90:1039-1044."""
    return check_n(lib().sfg_rnlm_len(_state["L"]))


def nlm_to_rnlm(nlm, rnlm_len):
    """Full -> reduced form, src/specfabpy.f90:1046-1054 -> reducedform.f90:176-183."""
    L = _L_of(nlm)
    if rnlm_len != check_n(lib().sfg_rnlm_len(L)):
        raise ValueError("rnlm_len=%d does not match len(nlm)=%d" % (rnlm_len, len(nlm)))
    r = np.zeros(rnlm_len, dtype=np.complex128)
    check(lib().sfg_nlm_to_rnlm(_ctx(), L, _zp(_z(nlm)), _zp(r)))
    return r


def rnlm_to_nlm(rnlm, nlm_len):
    """Reduced -> full form, src/specfabpy.f90:1056-1064 -> reducedform.f90:160-174."""
    L = _L_of_rnlm_len(len(rnlm))
    if nlm_len != (L + 1) * (L + 2) // 2:
        raise ValueError("nlm_len=%d does not match len(rnlm)=%d" % (nlm_len, len(rnlm)))
    nlm = np.zeros(nlm_len, dtype=np.complex128)
    check(lib().sfg_rnlm_to_nlm(_ctx(), L, _zp(_z(rnlm)), _zp(nlm)))
    return nlm


def reduce_M(M, rnlm_len):
    """(Mrr, Mri, Mir, Mii) of a full operator M, src/specfabpy.f90:1066-1074 ->
    reducedform.f90:76-120: d(rnlm)/dt = Mrr Re r + Mri Im r + i (Mir Re r + Mii Im r)."""
    M = np.ascontiguousarray(np.asarray(M, dtype=np.complex128))
    L = _L_of(M[0])
    if M.shape != (len(M[0]), len(M[0])) or rnlm_len != check_n(lib().sfg_rnlm_len(L)):
        raise ValueError("M of shape %s does not match rnlm_len=%d" % (M.shape, rnlm_len))
    out = [np.zeros((rnlm_len, rnlm_len)) for _ in range(4)]
    check(lib().sfg_reduce_M(_ctx(), L, _zp(M), *[dptr(o) for o in out]))
    return tuple(out)


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


def deformability_tranisotropic(nlm, tau, n_RSS):
    """<D> = ev_D(nlm, tau, n_RSS), the normalised average basal RSS^n_RSS
    (n_RSS = 2 or 4): src/specfabpy.f90:557-565 -> dynamics.f90:385-449."""
    L = _L_of(nlm)
    out = np.zeros(1)
    z = _z(nlm)
    check(lib().sfg_ev_D(_ctx(), L, _zp(z), dptr(_m3(tau)), int(n_RSS), dptr(out)))
    return float(out[0])


def deformability_tranisotropic_arr(nlm, tau, n_RSS):
    """<D> of every row of nlm (N, n) for one tau: src/specfabpy.f90:567-578,
    evaluated in one GPU launch."""
    import torch
    from .batch import FabricBatch
    nlm = np.asarray(nlm)
    b = FabricBatch(_L_of(nlm[0]), nlm=nlm)
    D = b.deformability(np.tile(_m3(tau), (len(nlm), 1, 1)), n_RSS)
    torch.cuda.synchronize()
    return D.cpu().numpy()


def E_CAFFE(nlm, eps, Emin, Emax, n_grain):
    """CAFFE enhancement factor (Placidi et al., 2010) of the strain rate eps:
    src/specfabpy.f90:461-470 -> enhancementfactors.f90:302-332."""
    L = _L_of(nlm)
    out = np.zeros(1)
    z = _z(nlm)
    check(lib().sfg_E_CAFFE(_ctx(), L, _zp(z), dptr(_m3(eps)), float(Emin), float(Emax), int(n_grain), dptr(out)))
    return float(out[0])


def E_CAFFE_arr(nlm, eps, Emin, Emax, n_grain):
    """E_CAFFE of nlm (N, n) with strain rates eps (N, 3, 3):
    src/specfabpy.f90:543-555, evaluated in one GPU launch."""
    import torch
    from .batch import FabricBatch
    nlm = np.asarray(nlm)
    b = FabricBatch(_L_of(nlm[0]), nlm=nlm)
    E = b.E_CAFFE(eps, Emin, Emax, n_grain)
    torch.cuda.synchronize()
    return E.cpu().numpy()


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


def Evw_tranisotropic(nlm, v, w, tau, Eij_grain, alpha, n_grain):
    """Enhancement factor of the stress state tau along v (x) w:
    src/specfabpy.f90:379-388 -> src/enhancementfactors.f90:47-69."""
    L = _L_of(nlm)
    f = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    out = np.zeros(1)
    check(lib().sfg_Evw_tranisotropic(_ctx(), L, _zp(_z(nlm)), dptr(f(v)), dptr(f(w)), dptr(_m3(tau)),
                                      dptr(f(Eij_grain).reshape(2)), float(alpha), int(n_grain), dptr(out)))
    return float(out[0])


def _mono_tensor(m, d):
    """the full symmetric (3,)*d tensor of the distinct moments m (monomials
    <x^a y^b z^c> in the order a = d..0, b = d - a..0)"""
    pos = {}
    k = 0
    for a in range(d, -1, -1):
        for b in range(d - a, -1, -1):
            pos[(a, b)] = k
            k += 1
    out = np.empty((3,) * d)
    for idx in np.ndindex(*(3,) * d):
        out[idx] = m[pos[(idx.count(0), idx.count(1))]]
    return out


def _moments68(nlm):
    L = _L_of(nlm)
    m6, m8 = np.zeros(28), np.zeros(45)
    check(lib().sfg_moments68(_ctx(), L, _zp(_z(nlm)), dptr(m6), dptr(m8)))
    return m6, m8


def a6(nlm):
    """Sixth-order structure tensor <c^6> (3,)*6: src/specfabpy.f90:601-608 ->
    src/moments.f90:57-66 (exact moments; the reference's body carries REAL(4)
    literals, ~3e-8 apart).  L >= 6."""
    return _mono_tensor(_moments68(nlm)[0], 6)


def ai(nlm):
    """(a2, a4, a6, a8) of nlm: src/specfabpy.f90:610-618 (a2 and a4 the
    reference's bodies, a6 and a8 exact).  L >= 8."""
    if _L_of(nlm) < 8:
        raise ValueError("ai: the eighth moments need L >= 8")
    m6, m8 = _moments68(nlm)
    return a2(nlm), a4(nlm), _mono_tensor(m6, 6), _mono_tensor(m8, 8)


def eig3(M):
    """eig3(M) -> (e1, e2, e3, lami) of a symmetric 3x3 M, largest eigenvalue
    first: src/specfabpy.f90:346-355 -> src/frames.f90:62-80 (on the GPU)."""
    ei = np.zeros(9)
    lam = np.zeros(3)
    check(lib().sfg_eig3(_ctx(), dptr(_m3(M)), dptr(ei), dptr(lam)))
    e = ei.reshape(3, 3)
    return e[0].copy(), e[1].copy(), e[2].copy(), lam


def _plane_sort(ei, lam, plane):
    """eigframe's ordering (src/frames.f90:24-60): 'ij' as eig3; 'xy' / 'xz'
    put the eigenvector most aligned with z / y last"""
    if plane == "ij":
        return ei, lam
    if plane not in ("xy", "xz"):
        raise ValueError('eigframe(M, plane): plane must be one of "ij", "xy", "xz"')
    k = 2 if plane == "xy" else 1
    imax = np.argmax(np.abs(ei[..., :, k]), axis=-1)
    order = np.array([[1, 2, 0], [0, 2, 1], [0, 1, 2]])[imax]
    return (np.take_along_axis(ei, order[..., :, None], axis=-2), np.take_along_axis(lam, order, axis=-1))


def eigframe(M, plane):
    """eigframe(M, plane) -> (ei (3, 3) rows, lami): src/specfabpy.f90:322-331."""
    e1, e2, e3, lam = eig3(M)
    return _plane_sort(np.stack([e1, e2, e3]), lam, plane)


def eigframe_arr(M, plane):
    """eigframe of every M (N, 3, 3) in one GPU launch: src/specfabpy.f90:333-344."""
    import torch
    Mt = torch.as_tensor(np.ascontiguousarray(np.asarray(M, dtype=np.float64).reshape(-1, 9)),
                         device=torch.device("cuda", torch.cuda.current_device()))
    N = Mt.shape[0]
    ei = torch.empty((N, 9), dtype=torch.float64, device=Mt.device)
    lam = torch.empty((N, 3), dtype=torch.float64, device=Mt.device)
    check(lib().sfg_eig3_batch(_ctx(), N, Mt.data_ptr(), ei.data_ptr(), lam.data_ptr(),
                               torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return _plane_sort(ei.cpu().numpy().reshape(N, 3, 3), lam.cpu().numpy(), plane)


def evD(nlm, tau):
    """<D> = ev_D2(nlm, tau): src/specfabpy.f90:1105-1111 (on the GPU)."""
    return deformability_tranisotropic(nlm, tau, 2)


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


def rotate_nlm(nlm, theta, phi):
    """rotate_nlm(nlm, theta, phi) (src/rotation.f90:29-55, specfabpy
    src/specfabpy.f90): n'_l^m = sum_n exp(-i n phi) d^l_mn(theta) n_l^n, L <= 12."""
    L = _L_of(nlm)
    z = _z(nlm)
    out = np.zeros_like(z)
    check(lib().sfg_rotate_nlm(_ctx(), L, _zp(z), float(theta), float(phi), _zp(out)))
    return out


def a4_eigentensors(nlm):
    """(Q1, ..., Q6, eigvals) of a4 (src/frames.f90:131-152): the Mandel
    matrix's eigenvalues (ascending) and eigentensors (3, 3)."""
    L = _L_of(nlm)
    z = _z(nlm)
    Q = np.zeros(54)
    ev = np.zeros(6)
    check(lib().sfg_a4_eigentensors(_ctx(), L, _zp(z), dptr(Q), dptr(ev)))
    Q = Q.reshape(6, 3, 3)
    return (*Q, ev)


def nlm_isvalid(nhat20, nhat40):
    """nlm_isvalid(nhat20, nhat40) (src/idealstate.f90:50-78): whether the
    states n_0^0 = 1, n_2^0 = nhat20, n_4^0 = nhat40 have a2 and a4
    eigenvalues inside (0, 1); array-valued like the reference."""
    n20 = np.ascontiguousarray(np.atleast_1d(nhat20), dtype=np.complex128)
    n40 = np.ascontiguousarray(np.atleast_1d(nhat40), dtype=np.complex128)
    if n20.shape != n40.shape:
        raise ValueError("nhat20 and nhat40 must have the same shape")
    ok = np.zeros(n20.size, dtype=np.int32)
    check(lib().sfg_nlm_isvalid(_ctx(), n20.size, _zp(n20.ravel()), _zp(n40.ravel()),
                                ok.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
    return ok.astype(bool).reshape(n20.shape)


# ---- discrete grains ----

def ri_LROT(ri0, dt, Nt, eps, omg, iota):
    """ri (Nt, N, 3): Euler steps of the grain axes ri0 (N, 3) under the
    discrete lattice rotation with strain rates eps (Nt, 3, 3) and spins
    omg (Nt, 3, 3), src/specfabpy.f90:271-280 -> dynamics.f90:112-137 (on
    the GPU)."""
    r0 = np.ascontiguousarray(np.asarray(ri0, dtype=np.float64))
    if r0.ndim != 2 or r0.shape[1] != 3:
        raise ValueError("ri0 must be (N, 3)")
    Nt = int(Nt)
    if Nt < 1:
        raise ValueError("Nt must be >= 1")
    D = np.ascontiguousarray(np.asarray(eps, dtype=np.float64))
    W = np.ascontiguousarray(np.asarray(omg, dtype=np.float64))
    if D.ndim != 3 or D.shape[1:] != (3, 3) or D.shape[0] < Nt - 1 or W.shape != D.shape:
        raise ValueError("eps and omg must be (>= Nt - 1, 3, 3) and alike")
    D = np.ascontiguousarray(D[:Nt]) if D.shape[0] >= Nt else np.vstack([D, np.zeros((1, 3, 3))])
    W = np.ascontiguousarray(W[:Nt]) if W.shape[0] >= Nt else np.vstack([W, np.zeros((1, 3, 3))])
    out = np.zeros((Nt, r0.shape[0], 3))
    check(lib().sfg_ri_LROT(_ctx(), r0.shape[0], Nt, dptr(r0), float(dt), dptr(D), dptr(W), float(iota),
                            dptr(out)))
    return out


def ri_to_nlm(ri, wi, L):
    """nlm ((L+1)(L+2)/2,) of the weighted grain-axis ensemble ri (N, 3),
    wi (N,), from its a2 (L = 2), a4 (L = 4) or a6 (L >= 6) moment,
    src/specfabpy.f90:676-684 -> moments.f90:98-131 (on the GPU)."""
    r = np.ascontiguousarray(np.asarray(ri, dtype=np.float64))
    w = np.ascontiguousarray(np.asarray(wi, dtype=np.float64).ravel())
    if r.ndim != 2 or r.shape[1] != 3 or w.size != r.shape[0]:
        raise ValueError("ri must be (N, 3) and wi (N,)")
    n = check_n(lib().sfg_nlm_len(int(L)))
    out = np.zeros(n, dtype=np.complex128)
    check(lib().sfg_ri_to_nlm(_ctx(), int(L), r.shape[0], dptr(r), dptr(w), _zp(out)))
    return out


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


# ---- orthotropic grains (olivine) ----

def Cij_to_Lame_orthotropic(Cij):
    """(C11,C22,C33,C44,C55,C66,C23,C13,C12) -> (lam11,lam22,lam33,lam23,lam13,lam12,mu1,mu2,mu3),
    src/elasticities.f90:122-139."""
    C = np.ascontiguousarray(np.asarray(Cij, dtype=np.float64).reshape(9))
    out = np.zeros(9)
    check(lib().sfg_Cij_to_Lame_orthotropic(dptr(C), dptr(out)))
    return out


def Lame_olivine_A2X(Lame, Xtype):
    """A-type olivine Lame parameters reordered for B- or C-type fabrics,
    src/elasticities.f90:141-154 (host arithmetic only)."""
    L = np.asarray(Lame, dtype=np.float64).reshape(9)
    order = {"B": [2, 1, 0, 5, 4, 3, 8, 7, 6], "C": [2, 0, 1, 5, 3, 4, 8, 6, 7]}
    if Xtype not in order:
        raise ValueError("Lame_olivine_A2X(): fabric type not supported")
    return L[order[Xtype]].copy()


def _ortho_states(q1, q2, q3):
    L = _L_of(q1)
    z1, z2 = _z(q1), _z(q2)
    z3 = _z(q3) if q3 is not None else np.zeros_like(z1)
    if len(z2) != len(z1) or len(z3) != len(z1):
        raise ValueError("the three distributions must have the same length")
    return L, z1, z2, z3


def Eij_orthotropic(nlm_r1, nlm_r2, nlm_r3, e1, e2, e3, Eij_grain, alpha, n_grain):
    """Enhancement factors of orthotropic grains (Sachs) w.r.t. (e1,e2,e3) from
    the distributions of the slip-system axes (b, n, v); nlm_r3 = 0 derives v =
    b x n: src/specfabpy.f90:436-446 -> enhancementfactors.f90:134-188."""
    L, z1, z2, z3 = _ortho_states(nlm_r1, nlm_r2, nlm_r3)
    v = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    out = np.zeros(6)
    check(lib().sfg_Eij_orthotropic(_ctx(), L, _zp(z1), _zp(z2), _zp(z3), dptr(v(e1)), dptr(v(e2)), dptr(v(e3)),
                                    dptr(v(Eij_grain).reshape(6)), float(alpha), int(n_grain), dptr(out)))
    return out


def Evw_orthotropic(nlm_r1, nlm_r2, nlm_r3, v, w, tau, Eij_grain, alpha, n_grain):
    """Orthotropic enhancement factor of the stress state tau along v (x) w:
    src/specfabpy.f90:425-434 -> src/enhancementfactors.f90:134-160."""
    L, z1, z2, z3 = _ortho_states(nlm_r1, nlm_r2, nlm_r3)
    f = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    out = np.zeros(1)
    check(lib().sfg_Evw_orthotropic(_ctx(), L, _zp(z1), _zp(z2), _zp(z3), dptr(f(v)), dptr(f(w)), dptr(_m3(tau)),
                                    dptr(f(Eij_grain).reshape(6)), float(alpha), int(n_grain), dptr(out)))
    return float(out[0])


def Eij_orthotropic_arr(nlm_r1, nlm_r2, nlm_r3, e1, e2, e3, Eij_grain, alpha, n_grain):
    """Batched orthotropic Eij (src/specfabpy.f90:488-500): nlm_ri (N, n), e_i
    (N, 3) -> (N, 6), one GPU launch."""
    import torch
    from .batch import FabricBatch, eij_orthotropic
    q1, q2 = np.asarray(nlm_r1), np.asarray(nlm_r2)
    L = _L_of(q1[0])
    b1, b2 = FabricBatch(L, nlm=q1), FabricBatch(L, nlm=q2)
    b3 = FabricBatch(L, nlm=np.asarray(nlm_r3)) if nlm_r3 is not None else None
    frames = np.stack([np.asarray(e1), np.asarray(e2), np.asarray(e3)], axis=1)
    E = eij_orthotropic(b1, b2, b3, Eij_grain, alpha=alpha, n_grain=n_grain, frames=frames)
    torch.cuda.synchronize()
    return E.cpu().numpy()


def Vi_elastic_orthotropic(nlm_1, nlm_2, nlm_3, alpha, lame_grain, rho, theta, phi):
    """Elastic phase velocities (qS1, qS2, qP) x ndir of orthotropic grains (Voigt),
    src/specfabpy.f90:975-984 -> wavepropagation.f90:41-68.  Returns (3, ndir)."""
    L, z1, z2, z3 = _ortho_states(nlm_1, nlm_2, nlm_3)
    th, ph = _dirs(theta, phi)
    out = np.zeros((th.size, 3))
    Lm = np.ascontiguousarray(np.asarray(lame_grain, dtype=np.float64).reshape(9))
    check(lib().sfg_Vi_elastic_orthotropic(_ctx(), L, _zp(z1), _zp(z2), _zp(z3), float(alpha), dptr(Lm), float(rho),
                                           th.size, dptr(th), dptr(ph), dptr(out)))
    return out.T.copy()


def Eij_orthotropic_discrete(bi, ni, vi, e1, e2, e3, Eij_grain, alpha, n_grain):
    """Eij of N discrete grains with slip-system axes bi, ni, vi (N, 3) w.r.t.
    (e1, e2, e3): src/enhancementfactors.f90:191-266 (Sachs)."""
    f = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    b, n, v = f(bi), f(ni), f(vi)
    if b.ndim != 2 or b.shape[1] != 3 or n.shape != b.shape or v.shape != b.shape:
        raise ValueError("bi, ni, vi must be (N, 3) each")
    out = np.zeros(6)
    check(lib().sfg_Eij_orthotropic_discrete(_ctx(), b.shape[0], dptr(b), dptr(n), dptr(v), dptr(f(e1)), dptr(f(e2)),
                                             dptr(f(e3)), dptr(f(Eij_grain).reshape(6)), float(alpha), int(n_grain),
                                             dptr(out)))
    return out


def Eij_orthotropic_discrete_arr(bi, ni, vi, e1, e2, e3, Eij_grain, alpha, n_grain):
    """Eij (Nt, 6) of Nt grain ensembles bi, ni, vi (Nt, N, 3) in frames e_i
    (Nt, 3) (src/specfabpy.f90:502-514), one GPU launch."""
    import torch
    from .batch import eij_orthotropic_discrete
    dev = torch.device("cuda", torch.cuda.current_device())
    t = lambda a: torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev)
    frames = np.stack([np.asarray(e1), np.asarray(e2), np.asarray(e3)], axis=1)
    E = eij_orthotropic_discrete(t(bi), t(ni), t(vi), Eij_grain, alpha=alpha, n_grain=n_grain, frames=t(frames))
    torch.cuda.synchronize()
    return E.cpu().numpy()


def Vi_elastic_orthotropic__discrete(mi, alpha, lame_grain, rho, theta, phi):
    """Elastic phase velocities (qS1, qS2, qP) x ndir of N discrete grains,
    mi (3, 3, N) = (axis, xyz, grain) (Voigt): src/specfabpy.f90:986-995 ->
    src/wavepropagation.f90:70-96.  Returns (3, ndir)."""
    m = np.asarray(mi, dtype=np.float64)
    if m.ndim != 3 or m.shape[:2] != (3, 3):
        raise ValueError("mi must be (3, 3, N)")
    b, n, v = (np.ascontiguousarray(m[i].T) for i in range(3))
    th, ph = _dirs(theta, phi)
    out = np.zeros((th.size, 3))
    Lm = np.ascontiguousarray(np.asarray(lame_grain, dtype=np.float64).reshape(9))
    check(lib().sfg_Vi_elastic_orthotropic_discrete(_ctx(), b.shape[0], dptr(b), dptr(n), dptr(v), float(alpha),
                                                    dptr(Lm), float(rho), th.size, dptr(th), dptr(ph), dptr(out)))
    return out.T.copy()


def Qnorm_tranisotropic(nlm, alpha, lame_grain):
    """(1 - alpha) Reuss + alpha Voigt acoustic tensor (3, 3) for k = z:
    src/specfabpy.f90:1010-1018 -> src/wavepropagation.f90:146-211."""
    L = _L_of(nlm)
    out = np.zeros(9)
    Lm = np.ascontiguousarray(np.asarray(lame_grain, dtype=np.float64).reshape(5))
    check(lib().sfg_Qnorm_tranisotropic(_ctx(), L, _zp(_z(nlm)), float(alpha), dptr(Lm), dptr(out)))
    return out.reshape(3, 3)


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


# ---- initial states, rate factor, bounds ----

def a2_to_nlm(a2):
    """nlm (6,) of the l <= 2 coefficients with structure tensor a2,
    src/specfabpy.f90:619-627 -> moments.f90:68-74 (on the GPU)."""
    out = np.zeros(6, dtype=np.complex128)
    check(lib().sfg_a2_to_nlm(_ctx(), dptr(_m3(a2)), _zp(out)))
    return out


def a4_to_nlm(a4):
    """nlm (15,) of the l <= 4 coefficients with structure tensor a4,
    src/specfabpy.f90:629-637 -> moments.f90:76-84 (on the GPU)."""
    a = np.ascontiguousarray(np.asarray(a4, dtype=np.float64).reshape(81))
    out = np.zeros(15, dtype=np.complex128)
    check(lib().sfg_a4_to_nlm(_ctx(), dptr(a), _zp(out)))
    return out


def Gamma0(eps, T, A, Q):
    """DDRX rate factor A sqrt(tr(eps.eps)/2) exp(-Q/(R T)),
    src/specfabpy.f90:1114-1122 -> dynamics.f90:451-472 (on the GPU)."""
    out = np.zeros(1)
    check(lib().sfg_Gamma0(_ctx(), dptr(_m3(eps)), float(T), float(A), float(Q), dptr(out)))
    return float(out[0])


def apply_bounds(nlm):
    """nlm with its l = 2 and l = 4 power spectrum bounded by that of a delta
    function, src/specfabpy.f90:764-771 -> dynamics.f90:530-557 (on the GPU)."""
    L = _L_of(nlm)
    z = _z(nlm)
    out = np.zeros(len(z), dtype=np.complex128)
    check(lib().sfg_apply_bounds(_ctx(), L, _zp(z), _zp(out)))
    return out


def _l_of_index(n):
    L = int(round((-3 + np.sqrt(1 + 8 * n)) / 2))
    return np.concatenate([np.full(2 * l + 1, l) for l in range(0, L + 1, 2)])


def Lmat(nlm):
    """Laplacian diag(-l(l+1)) (n, n), src/specfabpy.f90:247-258."""
    ls = _l_of_index(len(nlm))
    return np.diag(-(ls * (ls + 1)).astype(np.float64))


def Sl(nlm, l):
    """Angular power S(l) = sum_m |n_l^m|^2 / (2l + 1), src/idealstate.f90:80-93."""
    z = _z(nlm)
    ls = _l_of_index(len(z))
    return float(np.sum(np.abs(z[ls == l]) ** 2) / (2 * l + 1))


# ---- idealized kinematics (src/deformationmodes.f90) ----

def F_to_strain(F):
    F = np.asarray(F, dtype=np.float64)
    return 0.5 * (F + F.T) - np.eye(3)


def pureshear_r(tau, time):
    return np.exp(-time / tau)


def pureshear_strainii_to_t(strainii, tau):
    return -tau * np.log(strainii + 1.0)


def pureshear_F(ax, q, tau, time):
    r = pureshear_r(tau, time)
    qp, qm = -(1 + q) / 2.0, -(1 - q) / 2.0
    d = {0: (r, r ** qp, r ** qm), 1: (r ** qm, r, r ** qp), 2: (r ** qp, r ** qm, r)}[int(ax)]
    return np.diag(d)


def pureshear_ugrad(ax, q, tau):
    qp, qm = -(1 + q) / 2.0, -(1 - q) / 2.0
    d = {0: (1.0, qp, qm), 1: (qm, 1.0, qp), 2: (qp, qm, 1.0)}[int(ax)]
    return -1 / tau * np.diag(d)


def simpleshear_gamma(tau, time):
    return np.arctan(time / tau)


def simpleshear_gamma_to_t(gam, tau):
    return tau * np.tan(gam)


_SS_IJ = {0: (1, 2), 1: (0, 2), 2: (0, 1)}   # yz, xz, xy shear (du_i/dx_j)


def simpleshear_F(plane, tau, time):
    F = np.eye(3)
    F[_SS_IJ[int(plane)]] = np.tan(simpleshear_gamma(tau, time))
    return F


def simpleshear_ugrad(plane, tau):
    G = np.zeros((3, 3))
    G[_SS_IJ[int(plane)]] = 1 / tau
    return G


def __getattr__(name):
    # Lcap: the active truncation (src/specfabpy.f90:132,156)
    if name == "Lcap":
        return _state["L"]
    raise AttributeError(name)
