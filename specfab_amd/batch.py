"""Batched, device-resident interface: many Lagrangian parcels / Eulerian
nodes stepped and homogenized on one GPU.

PyTorch is used only for device memory and streams; every computation is a
call into libspecfab_gpu (HIP).  A batch holds the real SoA state x (n, P) in
HBM; see include/specfab_gpu.h for the layout.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import EvolveDesc, check, lib

_c = ctypes


def _ptr(t):
    return _c.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return _c.c_void_p(torch.cuda.current_stream().cuda_stream)


def nlm_len(L):
    return (L + 1) * (L + 2) // 2


def gamma0(eps, T, A, Q, ctx=None):
    """DDRX rate factors (P,) A sqrt(tr(eps.eps)/2) exp(-Q/(R T)) of strain
    rates eps (P, 3, 3) (device tensor) and temperatures T ((P,) tensor or a
    number), src/dynamics.f90:451-472."""
    ctx = ctx or _lib.default_context()
    E = eps.to(torch.float64).contiguous()
    P = E.shape[0]
    G = torch.empty(P, dtype=torch.float64, device=E.device)
    Tt = T.to(E.device, torch.float64).contiguous() if torch.is_tensor(T) else None
    check(lib().sfg_Gamma0_batch(ctx.handle, P, _ptr(E), _ptr(Tt), float(0.0 if Tt is not None else T), float(A),
                                 float(Q), _ptr(G), _stream()))
    return G


def rnlm_len(L):
    return (L + 2) ** 2 // 4


def _cplx(t):
    """complex128 device tensor, contiguous (its float64 view is the re/im interleaving the kernels read)"""
    return t.to(torch.complex128).contiguous()


def M_LROT_batch(L, eps, omg, iota, zeta, ctx=None):
    """Lattice-rotation operators (P, n, n) complex of P Eulerian nodes from
    strain rates eps and spins omg (P, 3, 3) (device tensors):
    src/specfabpy/fenics/CPO.py:200 over the whole mesh, src/dynamics.f90:52-97."""
    ctx = ctx or _lib.default_context()
    E, W = eps.to(torch.float64).contiguous(), omg.to(torch.float64).contiguous()
    n = nlm_len(L)
    M = torch.empty((E.shape[0], n, n), dtype=torch.complex128, device=E.device)
    check(lib().sfg_M_LROT_batch(ctx.handle, L, E.shape[0], _ptr(E), _ptr(W), float(iota), float(zeta), _ptr(M),
                                 _stream()))
    return M


def M_DDRX_src_batch(L, tau, ctx=None):
    """DDRX source operators (P, n, n) complex of P nodes from stresses tau
    (P, 3, 3): src/specfabpy/fenics/CPO.py:201, src/dynamics.f90:277-298."""
    ctx = ctx or _lib.default_context()
    T = tau.to(torch.float64).contiguous()
    n = nlm_len(L)
    M = torch.empty((T.shape[0], n, n), dtype=torch.complex128, device=T.device)
    check(lib().sfg_M_DDRX_src_batch(ctx.handle, L, T.shape[0], _ptr(T), _ptr(M), _stream()))
    return M


def M_REG_batch(L, D, ctx=None):
    """Regularization operators (P, n, n) float64 of P nodes from strain
    rates D (P, 3, 3): src/specfabpy/fenics/CPO.py:202, src/dynamics.f90:494-518."""
    ctx = ctx or _lib.default_context()
    Dd = D.to(torch.float64).contiguous()
    n = nlm_len(L)
    M = torch.empty((Dd.shape[0], n, n), dtype=torch.float64, device=Dd.device)
    check(lib().sfg_M_REG_batch(ctx.handle, L, Dd.shape[0], _ptr(Dd), _ptr(M), _stream()))
    return M


def reduce_M_batch(L, M, ctx=None):
    """(Mrr, Mri, Mir, Mii), each (P, rn, rn) float64, of full operators M
    (P, n, n) (device tensor): reduce_M (src/reducedform.f90:76-120) per node."""
    ctx = ctx or _lib.default_context()
    Mc = _cplx(M)
    P, rn = Mc.shape[0], rnlm_len(L)
    if Mc.shape[1:] != (nlm_len(L), nlm_len(L)):
        raise ValueError("M has shape %s, expected (P, %d, %d)" % (tuple(Mc.shape), nlm_len(L), nlm_len(L)))
    out = torch.empty((4, P, rn, rn), dtype=torch.float64, device=Mc.device)
    check(lib().sfg_reduce_M_batch(ctx.handle, L, P, _ptr(Mc), _ptr(out[0]), _ptr(out[1]), _ptr(out[2]),
                                   _ptr(out[3]), _stream()))
    return out[0], out[1], out[2], out[3]


def M_LROT_reduced_batch(L, eps, omg, iota, zeta, ctx=None):
    """reduce_M(M_LROT) of P nodes, (Mrr, Mri, Mir, Mii) each (P, rn, rn),
    built without the full (n, n) operators (the same values as
    reduce_M_batch(L, M_LROT_batch(...)))."""
    ctx = ctx or _lib.default_context()
    E, W = eps.to(torch.float64).contiguous(), omg.to(torch.float64).contiguous()
    P, rn = E.shape[0], rnlm_len(L)
    out = torch.empty((4, P, rn, rn), dtype=torch.float64, device=E.device)
    check(lib().sfg_M_LROT_reduced_batch(ctx.handle, L, P, _ptr(E), _ptr(W), float(iota), float(zeta), _ptr(out[0]),
                                         _ptr(out[1]), _ptr(out[2]), _ptr(out[3]), _stream()))
    return out[0], out[1], out[2], out[3]


def M_DDRX_src_reduced_batch(L, tau, ctx=None):
    """reduce_M(M_DDRX_src) of P nodes without the full operators."""
    ctx = ctx or _lib.default_context()
    T = tau.to(torch.float64).contiguous()
    P, rn = T.shape[0], rnlm_len(L)
    out = torch.empty((4, P, rn, rn), dtype=torch.float64, device=T.device)
    check(lib().sfg_M_DDRX_src_reduced_batch(ctx.handle, L, P, _ptr(T), _ptr(out[0]), _ptr(out[1]), _ptr(out[2]),
                                             _ptr(out[3]), _stream()))
    return out[0], out[1], out[2], out[3]


def nlm_to_rnlm_batch(L, nlm, ctx=None):
    """(P, n) complex -> (P, rn) complex, src/reducedform.f90:176-183 per node."""
    ctx = ctx or _lib.default_context()
    z = _cplx(nlm)
    if z.shape[1] != nlm_len(L):
        raise ValueError("nlm has %d coefficients, expected %d" % (z.shape[1], nlm_len(L)))
    r = torch.empty((z.shape[0], rnlm_len(L)), dtype=torch.complex128, device=z.device)
    check(lib().sfg_nlm_to_rnlm_batch(ctx.handle, L, z.shape[0], _ptr(z), _ptr(r), _stream()))
    return r


def rnlm_to_nlm_batch(L, rnlm, ctx=None):
    """(P, rn) complex -> (P, n) complex, src/reducedform.f90:160-174 per node."""
    ctx = ctx or _lib.default_context()
    r = _cplx(rnlm)
    if r.shape[1] != rnlm_len(L):
        raise ValueError("rnlm has %d coefficients, expected %d" % (r.shape[1], rnlm_len(L)))
    z = torch.empty((r.shape[0], nlm_len(L)), dtype=torch.complex128, device=r.device)
    check(lib().sfg_rnlm_to_nlm_batch(ctx.handle, L, r.shape[0], _ptr(r), _ptr(z), _stream()))
    return z


def ri_LROT_batch(ri0, dt, Nt, eps, omg, iota, ctx=None):
    """Trajectories ri (Nt, N, 3) of N grain axes ri0 (N, 3) (device tensor)
    under the discrete lattice rotation with per-step strain rates eps and
    spins omg (Nt, 3, 3): ri_LROT, src/dynamics.f90:112-137."""
    ctx = ctx or _lib.default_context()
    r0 = ri0.to(torch.float64).contiguous()
    if r0.dim() != 2 or r0.shape[1] != 3:
        raise ValueError("ri0 must be (N, 3)")
    Nt = int(Nt)
    if Nt < 1:
        raise ValueError("Nt must be >= 1")
    D = eps.to(r0.device, torch.float64).contiguous()
    W = omg.to(r0.device, torch.float64).contiguous()
    if D.shape != (Nt, 3, 3) or W.shape != (Nt, 3, 3):
        raise ValueError("eps and omg must be (Nt, 3, 3)")
    ri = torch.empty((Nt, r0.shape[0], 3), dtype=torch.float64, device=r0.device)
    check(lib().sfg_ri_LROT_batch(ctx.handle, r0.shape[0], Nt, _ptr(r0), float(dt), _ptr(D), _ptr(W), float(iota),
                                  _ptr(ri), _stream()))
    return ri


def eij_orthotropic(b1, b2, b3, Eij_grain, alpha=0.0, n_grain=1, frames=None, ctx=None):
    """Orthotropic Eij (P, 6) (Sachs; src/enhancementfactors.f90:134-188) of the
    slip-system-axis distributions held by FabricBatch objects b1 (b), b2 (n)
    and b3 (v; None or zero states derive v = b x n), in the caller frames
    (P, 3, 3) or the Cartesian frame (frames None)."""
    ctx = ctx or b1.ctx
    if b2.P != b1.P or b2.L != b1.L or (b3 is not None and (b3.P != b1.P or b3.L != b1.L)):
        raise ValueError("the three batches must have the same P and L")
    fr = b1._f64(frames, (b1.P, 3, 3)) if frames is not None else None
    E = torch.empty((b1.P, 6), dtype=torch.float64, device=b1.dev)
    Eg = (_c.c_double * 6)(*np.asarray(Eij_grain, dtype=np.float64).reshape(6))
    check(lib().sfg_Eij_orthotropic_batch(ctx.handle, b1.L, b1.P, _ptr(b1.x), _ptr(b2.x),
                                          _ptr(b3.x) if b3 is not None else None, 1 if fr is not None else 2,
                                          _ptr(fr), Eg, float(alpha), int(n_grain), _ptr(E), _stream()))
    return E


def _states(v, w, tau, dev):
    v = np.asarray(v, dtype=np.float64).reshape(-1, 3)
    w = np.asarray(w, dtype=np.float64).reshape(-1, 3)
    tau = np.asarray(tau, dtype=np.float64).reshape(-1, 3, 3)
    if not (len(v) == len(w) == len(tau)):
        raise ValueError("v, w, tau must hold the same number of stress states")
    if np.abs(tau - np.transpose(tau, (0, 2, 1))).max(initial=0) > 1e-12 * np.abs(tau).max(initial=0):
        raise ValueError("tau must be symmetric (a stress)")
    st = np.concatenate([v, w, tau.reshape(-1, 9)], axis=1)
    return torch.as_tensor(np.ascontiguousarray(st), device=dev)


def evw_orthotropic(b1, b2, b3, v, w, tau, Eij_grain, alpha=0.0, n_grain=1, ctx=None):
    """Orthotropic Evw (P, S) of the S stress states tau (S, 3, 3) along
    v (x) w (S, 3) for every parcel (src/enhancementfactors.f90:134-160)."""
    ctx = ctx or b1.ctx
    st = _states(v, w, tau, b1.dev)
    E = torch.empty((b1.P, st.shape[0]), dtype=torch.float64, device=b1.dev)
    Eg = (_c.c_double * 6)(*np.asarray(Eij_grain, dtype=np.float64).reshape(6))
    check(lib().sfg_Evw_orthotropic_batch(ctx.handle, b1.L, b1.P, _ptr(b1.x), _ptr(b2.x),
                                          _ptr(b3.x) if b3 is not None else None, st.shape[0], _ptr(st), Eg,
                                          float(alpha), int(n_grain), _ptr(E), _stream()))
    return E


def vi_elastic_orthotropic(b1, b2, b3, lame_grain, rho, theta, phi, alpha=1.0, ctx=None):
    """Orthotropic elastic phase velocities (P, ndir, 3) = (qS1, qS2, qP)
    (Voigt; src/wavepropagation.f90:41-68) of the distributions in b1, b2, b3
    (b3 None or zero: v = b x n) along (theta, phi)."""
    ctx = ctx or b1.ctx
    th = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
    ph = np.ascontiguousarray(np.atleast_1d(phi), dtype=np.float64)
    Vi = torch.empty((b1.P, th.size, 3), dtype=torch.float64, device=b1.dev)
    Lm = (_c.c_double * 9)(*np.asarray(lame_grain, dtype=np.float64).reshape(9))
    check(lib().sfg_Vi_elastic_orthotropic_batch(ctx.handle, b1.L, b1.P, _ptr(b1.x), _ptr(b2.x),
                                                 _ptr(b3.x) if b3 is not None else None, float(alpha), Lm,
                                                 float(rho), th.size,
                                                 th.ctypes.data_as(_lib._dp), ph.ctypes.data_as(_lib._dp),
                                                 _ptr(Vi), _stream()))
    return Vi


def _axes(t, name):
    t = torch.as_tensor(t).to(torch.float64)
    if t.dim() == 2:
        t = t.unsqueeze(0)
    if t.dim() != 3 or t.shape[2] != 3:
        raise ValueError("%s must be (E, N, 3) or (N, 3)" % name)
    return t.contiguous()


def eij_orthotropic_discrete(bi, ni, vi, Eij_grain, alpha=0.0, n_grain=1, frames=None, ctx=None):
    """Orthotropic Eij (E, 6) (Sachs; src/enhancementfactors.f90:191-266) of E
    ensembles of N grains with slip-system axes bi, ni, vi (E, N, 3) device
    tensors, in the caller frames (E, 3, 3) or the Cartesian frame."""
    ctx = ctx or _lib.default_context()
    b, n, v = _axes(bi, "bi"), _axes(ni, "ni"), _axes(vi, "vi")
    if n.shape != b.shape or v.shape != b.shape:
        raise ValueError("bi, ni, vi must have the same shape")
    E, N = b.shape[0], b.shape[1]
    fr = None
    if frames is not None:
        fr = torch.as_tensor(frames).to(b.device, torch.float64).reshape(E, 3, 3).contiguous()
    out = torch.empty((E, 6), dtype=torch.float64, device=b.device)
    Eg = (ctypes.c_double * 6)(*np.asarray(Eij_grain, dtype=np.float64).reshape(6))
    check(lib().sfg_Eij_orthotropic_discrete_batch(ctx.handle, E, N, _ptr(b), _ptr(n), _ptr(v),
                                                   1 if fr is not None else 2, _ptr(fr) if fr is not None else None,
                                                   Eg, float(alpha), int(n_grain), _ptr(out), _stream()))
    return out


def vi_elastic_orthotropic_discrete(bi, ni, vi, lame_grain, rho, theta, phi, alpha=1.0, ctx=None):
    """Orthotropic elastic phase velocities (E, ndir, 3) = (qS1, qS2, qP)
    (Voigt; src/wavepropagation.f90:70-96) of E ensembles of N grains with
    axes bi, ni, vi (E, N, 3) along (theta, phi)."""
    ctx = ctx or _lib.default_context()
    b, n, v = _axes(bi, "bi"), _axes(ni, "ni"), _axes(vi, "vi")
    if n.shape != b.shape or v.shape != b.shape:
        raise ValueError("bi, ni, vi must have the same shape")
    E, N = b.shape[0], b.shape[1]
    th = np.ascontiguousarray(np.atleast_1d(theta), dtype=np.float64)
    ph = np.ascontiguousarray(np.atleast_1d(phi), dtype=np.float64)
    Vi = torch.empty((E, th.size, 3), dtype=torch.float64, device=b.device)
    Lm = (ctypes.c_double * 9)(*np.asarray(lame_grain, dtype=np.float64).reshape(9))
    check(lib().sfg_Vi_elastic_orthotropic_discrete_batch(ctx.handle, E, N, _ptr(b), _ptr(n), _ptr(v), float(alpha),
                                                          Lm, float(rho), th.size, th.ctypes.data_as(_lib._dp),
                                                          ph.ctypes.data_as(_lib._dp), _ptr(Vi), _stream()))
    return Vi


class FabricBatch:
    """P parcels at truncation L with per-parcel kinematics.

    ugrad: (P,3,3) velocity gradients (D = sym, W = skew); tau: optional
    (P,3,3) stresses (default tau := D); gam0 / lam: DDRX / CDRX rates (P,).
    """

    def __init__(self, L, nlm=None, P=None, ugrad=None, tau=None, gam0=None, lam=None, steps=None,
                 ctx=None, device="cuda:0"):
        if not torch.cuda.is_available():
            raise _lib.SFGError("no HIP device available")
        self.ctx = ctx or _lib.default_context()
        self.L = int(L)
        self.n = nlm_len(self.L)
        self.dev = torch.device(device)
        if nlm is not None:
            nlm_t = torch.as_tensor(np.ascontiguousarray(nlm) if isinstance(nlm, np.ndarray) else nlm)
            nlm_t = nlm_t.to(self.dev, torch.complex128).contiguous()
            self.P = nlm_t.shape[0]
            self.x = torch.empty((self.n, self.P), dtype=torch.float64, device=self.dev)
            self.set_nlm(nlm_t)
        else:
            self.P = int(P)
            self.x = torch.zeros((self.n, self.P), dtype=torch.float64, device=self.dev)
            self.x[0] = 1 / np.sqrt(4 * np.pi)
        self.ugrad = self._f64(ugrad, (self.P, 3, 3)) if ugrad is not None else None
        self.tau = self._f64(tau, (self.P, 3, 3)) if tau is not None else None
        self.gam0 = self._f64(gam0, (self.P,)) if gam0 is not None else None
        self.lam = self._f64(lam, (self.P,)) if lam is not None else None
        self.steps = (torch.as_tensor(steps).to(self.dev, torch.int32).contiguous() if steps is not None else None)
        self.status = torch.zeros(self.P, dtype=torch.int32, device=self.dev)

    def _f64(self, a, shape):
        t = torch.as_tensor(np.asarray(a) if not torch.is_tensor(a) else a)
        t = t.to(self.dev, torch.float64)
        t = t.expand(shape) if t.dim() < len(shape) else t
        return t.reshape(shape).contiguous()

    # ---- state conversion ----
    def set_nlm(self, nlm):
        nlm_t = torch.as_tensor(nlm).to(self.dev, torch.complex128).contiguous()
        assert nlm_t.shape == (self.P, self.n)
        check(lib().sfg_nlm_to_state(self.ctx.handle, self.L, self.P, _ptr(torch.view_as_real(nlm_t)), _ptr(self.x),
                                     _stream()))

    def nlm(self):
        out = torch.empty((self.P, self.n), dtype=torch.complex128, device=self.dev)
        check(lib().sfg_state_to_nlm(self.ctx.handle, self.L, self.P, _ptr(self.x), _ptr(torch.view_as_real(out)),
                                     _stream()))
        return out

    # ---- evolution ----
    def evolve(self, nsteps, dt, scheme="rk4", procs=("LROT", "REG"), iota=1.0, zeta=0.0, precision="fp64",
               reg=None):
        """Advance every parcel nsteps (or its own step count) steps.  reg =
        (nu, expo) replaces the calibrated regularization of L by
        -nu |D| (l(l+1) / (L(L+1)))^expo."""
        if self.ugrad is None:
            raise ValueError("ugrad is required")
        mask = 0
        for p in procs:
            mask |= _lib.PROC_BITS[p]
        d = EvolveDesc()
        d.L = self.L
        d.P = self.P
        d.procs = mask
        d.scheme = _lib.SFG_EULER if scheme == "euler" else _lib.SFG_RK4
        d.precision = _lib.SFG_FP64 if precision == "fp64" else _lib.SFG_FP32C
        d.nsteps = int(nsteps)
        d.dt = float(dt)
        d.iota = float(iota)
        d.zeta = float(zeta)
        d.ugrad_dev = self.ugrad.data_ptr()
        d.tau_dev = self.tau.data_ptr() if self.tau is not None else None
        d.gam0_dev = self.gam0.data_ptr() if self.gam0 is not None else None
        d.lam_dev = self.lam.data_ptr() if self.lam is not None else None
        d.steps_dev = self.steps.data_ptr() if self.steps is not None else None
        d.x_dev = self.x.data_ptr()
        d.status_dev = self.status.data_ptr()
        if reg is not None:
            d.reg_nu, d.reg_expo = float(reg[0]), float(reg[1])
        check(lib().sfg_evolve(self.ctx.handle, _c.byref(d), _stream()))
        return self

    # ---- initial states and bounds ----
    def set_a2(self, a2):
        """States whose l <= 2 part has the structure tensors a2 (P, 3, 3)
        (a2_to_nlm, src/moments.f90:68-74); higher coefficients zeroed."""
        A = self._f64(a2, (self.P, 3, 3))
        check(lib().sfg_a2_to_state_batch(self.ctx.handle, self.L, self.P, _ptr(A), _ptr(self.x), _stream()))
        return self

    def set_a4(self, a4):
        """States whose l <= 4 part has the structure tensors a4 (P, 3, 3, 3, 3)
        (a4_to_nlm, src/moments.f90:76-84); higher coefficients zeroed."""
        A = self._f64(a4, (self.P, 3, 3, 3, 3))
        check(lib().sfg_a4_to_state_batch(self.ctx.handle, self.L, self.P, _ptr(A), _ptr(self.x), _stream()))
        return self

    def set_grains(self, ri, wi):
        """States of P grain ensembles ri (P, N, 3) with the shared weights
        wi (N,) (ri_to_nlm, src/moments.f90:98-131): from the a2 (L = 2), a4
        (L = 4) or a6 moment (L >= 6; coefficients above l = 6 zeroed)."""
        R = self._f64(ri, (self.P, -1, 3))
        w = torch.as_tensor(wi, dtype=torch.float64, device=self.dev).contiguous()
        if w.dim() != 1 or w.shape[0] != R.shape[1]:
            raise ValueError("wi must be (N,) with N = ri.shape[1]")
        check(lib().sfg_ri_to_state_batch(self.ctx.handle, self.L, self.P, R.shape[1], _ptr(R), _ptr(w),
                                          _ptr(self.x), _stream()))
        return self

    def apply_bounds(self):
        """In place: l = 2, 4 power spectra bounded by a delta function's
        (apply_bounds, src/dynamics.f90:530-557)."""
        check(lib().sfg_apply_bounds_batch(self.ctx.handle, self.L, self.P, _ptr(self.x), _stream()))
        return self

    def autotune(self, precision="fp64", P=None, steps=4):
        """Time the generated stepper variants for this L and precision on a
        synthetic batch (default: min(P, 2^18) parcels) in launches of
        `steps` steps (the steps per evolve call the caller will make) and
        keep the fastest for the context.  Returns (chosen variant,
        per-variant ms per step)."""
        Pp = int(P or min(self.P, 1 << 18))
        best = _c.c_int(-1)
        ms = (_c.c_double * 64)()
        prec = _lib.SFG_FP64 if precision == "fp64" else _lib.SFG_FP32C
        check(lib().sfg_autotune(self.ctx.handle, self.L, prec, Pp, int(steps), _stream(), _c.byref(best), ms, 64))
        times = [ms[i] for i in range(64) if ms[i] > 0]
        return best.value, times

    # ---- moments ----
    def a2(self):
        out = torch.empty((self.P, 3, 3), dtype=torch.float64, device=self.dev)
        check(lib().sfg_a2_batch(self.ctx.handle, self.L, self.P, _ptr(self.x), _ptr(out), _stream()))
        return out

    def a4_mandel(self):
        out = torch.empty((self.P, 6, 6), dtype=torch.float64, device=self.dev)
        check(lib().sfg_a4mandel_batch(self.ctx.handle, self.L, self.P, _ptr(self.x), _ptr(out), _stream()))
        return out

    # ---- deformability, CAFFE ----
    def deformability(self, tau, n_RSS=2):
        """<D> (P,) of per-parcel stresses tau (P, 3, 3), n_RSS = 2 or 4:
        ev_D, src/dynamics.f90:385-449."""
        T = self._f64(tau, (self.P, 3, 3))
        D = torch.empty(self.P, dtype=torch.float64, device=self.dev)
        check(lib().sfg_ev_D_batch(self.ctx.handle, self.L, self.P, _ptr(self.x), _ptr(T), int(n_RSS), _ptr(D),
                                   _stream()))
        return D

    def E_CAFFE(self, eps, Emin, Emax, n_grain=1):
        """CAFFE enhancement factors (P,) of per-parcel strain rates eps
        (P, 3, 3): E_CAFFE, src/enhancementfactors.f90:302-332."""
        T = self._f64(eps, (self.P, 3, 3))
        E = torch.empty(self.P, dtype=torch.float64, device=self.dev)
        check(lib().sfg_E_CAFFE_batch(self.ctx.handle, self.L, self.P, _ptr(self.x), _ptr(T), float(Emin),
                                      float(Emax), int(n_grain), _ptr(E), _stream()))
        return E

    # ---- homogenization ----
    def Eij(self, Eij_grain=(1.0, 1e3), alpha=0.0125, n_grain=1, frame="eig", frames=None, status=None):
        """Enhancement factors (P, 6) of transversely isotropic grains
        (Eij_tranisotropic, src/enhancementfactors.f90:23-45; n_grain = 1, -3
        or 3, the last with L >= 8) in the a2
        eigenframe ("eig", returns also eigenvalues (P,3) and eigenvectors
        (P,3,3)), the Cartesian frame ("xyz"), or caller frames (P,3,3)."""
        mode = {"eig": 0, "frames": 1, "xyz": 2}[frame if frames is None else "frames"]
        fr = self._f64(frames, (self.P, 3, 3)) if frames is not None else None
        E = torch.empty((self.P, 6), dtype=torch.float64, device=self.dev)
        lam = torch.empty((self.P, 3), dtype=torch.float64, device=self.dev) if mode == 0 else None
        ei = torch.empty((self.P, 3, 3), dtype=torch.float64, device=self.dev) if mode == 0 else None
        st = status if status is not None else self.status
        Eg = (_c.c_double * 2)(*Eij_grain)
        check(lib().sfg_Eij_tranisotropic_batch(self.ctx.handle, self.L, self.P, _ptr(self.x), mode, _ptr(fr), Eg,
                                                float(alpha), int(n_grain), _ptr(E), _ptr(lam), _ptr(ei), _ptr(st),
                                                _stream()))
        return (E, lam, ei) if mode == 0 else E

    # ---- rotation, a4 eigentensors, validity ----
    def Evw(self, v, w, tau, Eij_grain=(1.0, 1e3), alpha=0.0125, n_grain=1, status=None):
        """Evw (P, S) of the S stress states tau (S, 3, 3) along v (x) w (S, 3)
        for every parcel (src/enhancementfactors.f90:47-69)."""
        st = _states(v, w, tau, self.dev)
        E = torch.empty((self.P, st.shape[0]), dtype=torch.float64, device=self.dev)
        Eg = (_c.c_double * 2)(*np.asarray(Eij_grain, dtype=np.float64).reshape(2))
        check(lib().sfg_Evw_tranisotropic_batch(self.ctx.handle, self.L, self.P, _ptr(self.x), st.shape[0], _ptr(st),
                                                Eg, float(alpha), int(n_grain), _ptr(E),
                                                _ptr(status if status is not None else self.status), _stream()))
        return E

    def rotate(self, theta, phi):
        """Rotate every parcel's state in place by its own (theta, phi) (P)
        or by scalars: rotate_nlm, src/rotation.f90:29-55 (L <= 12)."""
        th = torch.as_tensor(theta, dtype=torch.float64, device=self.dev).expand(self.P).contiguous()
        ph = torch.as_tensor(phi, dtype=torch.float64, device=self.dev).expand(self.P).contiguous()
        check(lib().sfg_rotate_nlm_batch(self.ctx.handle, self.L, self.P, _ptr(self.x), _ptr(th), _ptr(ph),
                                         _ptr(self.x), _stream()))

    def a4_eigentensors(self):
        """(Q (P, 6, 3, 3), eigenvalues (P, 6) ascending): a4_eigentensors,
        src/frames.f90:131-152."""
        Q = torch.empty((self.P, 6, 3, 3), dtype=torch.float64, device=self.dev)
        ev = torch.empty((self.P, 6), dtype=torch.float64, device=self.dev)
        check(lib().sfg_a4_eigentensors_batch(self.ctx.handle, self.L, self.P, _ptr(self.x), _ptr(Q), _ptr(ev),
                                              _stream()))
        return Q, ev

    def isvalid(self):
        """(P) bool: a2 and a4 eigenvalues inside (0, 1), the test of
        nlm_isvalid (src/idealstate.f90:50-78) on general states."""
        ok = torch.empty(self.P, dtype=torch.int32, device=self.dev)
        check(lib().sfg_state_isvalid_batch(self.ctx.handle, self.L, self.P, _ptr(self.x), _ptr(ok), _stream()))
        return ok.bool()

    # ---- full a4 and wave velocities ----
    def a4(self):
        """Full a4 (P, 3, 3, 3, 3) with the reference body's REAL(4) literals."""
        out = torch.empty((self.P, 3, 3, 3, 3), dtype=torch.float64, device=self.dev)
        check(lib().sfg_a4_batch(self.ctx.handle, self.L, self.P, _ptr(self.x), _ptr(out), _stream()))
        return out

    def Vi_elastic(self, alpha, Lame_grain, rho, theta, phi, status=None):
        """Elastic phase velocities (P, ndir, 3) = (qS1, qS2, qP) per direction
        (Vi_elastic_tranisotropic, src/wavepropagation.f90:121-144)."""
        th = np.ascontiguousarray(np.atleast_1d(np.asarray(theta, dtype=np.float64)).ravel())
        ph = np.ascontiguousarray(np.atleast_1d(np.asarray(phi, dtype=np.float64)).ravel())
        Lm = np.ascontiguousarray(np.asarray(Lame_grain, dtype=np.float64).reshape(5))
        out = torch.empty((self.P, th.size, 3), dtype=torch.float64, device=self.dev)
        st = status if status is not None else self.status
        check(lib().sfg_Vi_elastic_tranisotropic_batch(
            self.ctx.handle, self.L, self.P, _ptr(self.x), float(alpha), _lib.dptr(Lm), float(rho), th.size,
            _lib.dptr(th), _lib.dptr(ph), _ptr(out), _ptr(st), _stream()))
        return out

    def Vi_electromagnetic(self, eps_m, eps_t, mu, theta, phi):
        """EM phase velocities (P, ndir, 2) = (S1, S2)
        (Vi_electromagnetic_tranisotropic, src/wavepropagation.f90:217-261)."""
        th = np.ascontiguousarray(np.atleast_1d(np.asarray(theta, dtype=np.float64)).ravel())
        ph = np.ascontiguousarray(np.atleast_1d(np.asarray(phi, dtype=np.float64)).ravel())
        out = torch.empty((self.P, th.size, 2), dtype=torch.float64, device=self.dev)
        check(lib().sfg_Vi_electromagnetic_tranisotropic_batch(
            self.ctx.handle, self.L, self.P, _ptr(self.x), float(eps_m), float(eps_t), float(mu), th.size,
            _lib.dptr(th), _lib.dptr(ph), _ptr(out), _stream()))
        return out


class UnevenBatch:
    """Parcels of different truncations L and step counts,
    stepped through their whole time loop by one graph launch.

    Ls: (P,) truncations; nlm: sequence of P complex arrays (length
    nlm_len(L_p)); ugrad (P, 3, 3); gam0, lam (P,) or scalars; steps (P,)
    step counts.  Parcels are bucketed by L and sorted by step count inside a
    bucket (wavefronts of similar length); results come back in the input
    order."""

    def __init__(self, Ls, nlm, ugrad, gam0=0.0, lam=0.0, steps=None, ctx=None, device="cuda:0"):
        self.ctx = ctx or _lib.default_context()
        Ls = np.asarray(Ls, dtype=np.int64)
        P = Ls.size
        ug = np.asarray(ugrad, dtype=np.float64).reshape(P, 3, 3)
        g0 = np.broadcast_to(np.asarray(gam0, dtype=np.float64), (P,))
        lm = np.broadcast_to(np.asarray(lam, dtype=np.float64), (P,))
        st = np.asarray(steps, dtype=np.int64).reshape(P)
        self.P = P
        self.buckets = []   # (L, original indices, FabricBatch, remaining-steps tensor)
        for L in sorted(set(Ls.tolist())):
            idx = np.nonzero(Ls == L)[0]
            idx = idx[np.argsort(st[idx], kind="stable")]
            X = np.array([np.asarray(nlm[i], dtype=np.complex128) for i in idx])
            fb = FabricBatch(L, nlm=X, ugrad=ug[idx], gam0=g0[idx], lam=lm[idx], ctx=self.ctx, device=device)
            rem = torch.as_tensor(st[idx].astype(np.int32)).to(fb.dev).contiguous()
            self.buckets.append((L, idx, fb, rem))
        self.max_steps = int(st.max()) if P else 0
        self.plan = None

    def make_plan(self, dt, scheme="rk4", procs=("LROT", "REG", "DDRX", "CDRX"), iota=1.0, zeta=0.0,
                  precision="fp64", chunk=50):
        mask = 0
        for p in procs:
            mask |= _lib.PROC_BITS[p]
        descs = (EvolveDesc * len(self.buckets))()
        for d, (L, idx, fb, rem) in zip(descs, self.buckets):
            d.L, d.P, d.procs = L, fb.P, mask
            d.scheme = _lib.SFG_EULER if scheme == "euler" else _lib.SFG_RK4
            d.precision = _lib.SFG_FP64 if precision == "fp64" else _lib.SFG_FP32C
            d.nsteps, d.dt, d.iota, d.zeta = 0, float(dt), float(iota), float(zeta)
            d.ugrad_dev = fb.ugrad.data_ptr()
            d.tau_dev = None
            d.gam0_dev = fb.gam0.data_ptr() if fb.gam0 is not None else None
            d.lam_dev = fb.lam.data_ptr() if fb.lam is not None else None
            d.steps_dev = rem.data_ptr()
            d.x_dev = fb.x.data_ptr()
            d.status_dev = fb.status.data_ptr()
        nchunks = max(1, -(-self.max_steps // chunk))
        h = _c.c_void_p()
        check(lib().sfg_plan_create(self.ctx.handle, len(self.buckets), descs, int(chunk), nchunks, _stream(),
                                    _c.byref(h)))
        self.close()
        self.plan = h
        return self

    def run(self):
        """One graph launch: every parcel takes all its remaining steps."""
        check(lib().sfg_plan_run(self.plan, _stream()))
        return self

    def close(self):
        if self.plan:
            lib().sfg_plan_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nlm(self):
        out = [None] * self.P
        for (L, idx, fb, rem) in self.buckets:
            X = fb.nlm().cpu().numpy()
            for k, i in enumerate(idx):
                out[i] = X[k]
        return out

    def status(self):
        s = np.zeros(self.P, np.int32)
        for (L, idx, fb, rem) in self.buckets:
            s[idx] = fb.status.cpu().numpy()
        return s

    def remaining(self):
        r = np.zeros(self.P, np.int64)
        for (L, idx, fb, rem) in self.buckets:
            r[idx] = rem.cpu().numpy()
        return r
