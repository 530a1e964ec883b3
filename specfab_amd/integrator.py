"""Lagrangian parcel integrator with the signature and results of the
reference's `specfabpy.integrator.lagrangianparcel`
(src/specfabpy/integrator.py:6-82), stepping on the GPU.

    from specfab_amd import specfabpy as sf
    from specfab_amd.integrator import lagrangianparcel
    sf.init(8)
    DK = dict(type="ps", q=0, tau=1, axis=2)
    nlm, F, time, ugrad = lagrangianparcel(sf, DK, -0.9, Nt=100, iota=1, nu=1)

The reference builds the steady operators (M_LROT, Lambda M_CDRX, nu M_REG or
the custom regularization) once, rebuilds Gamma0 M_DDRX(nlm, S) from the
previous state every step, advances by explicit Euler with dense complex
matrix-vector products and keeps the trajectory in complex64.  Here each step
is one launch of the batched FP64 Euler stepper (sfg_evolve, the same
operators in the real basis, S := D as in the reference) on a one-parcel
batch, and the state is rounded to complex64 after every step as the
reference's storage does, so the trajectories agree to float32 rounding.
`apply_bounds` is accepted and, as in the reference, not used.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .batch import FabricBatch


def _reg_law(sf, L, nu, regexpo):
    """(nu_eff, expo) of the regularization: nu x the calibrated M_REG
    (src/specfabpy/integrator.py:55), or the custom law nu |D|
    |l(l+1)/(L(L+1))|^regexpo (src/specfabpy/integrator.py:56, 85-97)."""
    if regexpo is not None:
        return float(nu), float(regexpo)
    expo, nu_cal = ctypes.c_double(), ctypes.c_double()
    _lib.check(_lib.lib().sfg_reg_calibration(int(L), ctypes.byref(expo), ctypes.byref(nu_cal)))
    return float(nu) * nu_cal.value, expo.value


def lagrangianparcel(sf, DK, strain_target, Nt=100, dt=None, nlm0=None, verbose=True,
                     iota=None, zeta=0, Lambda=None, Gamma0=None,
                     nu=None, regexpo=None, apply_bounds=False):
    """Lagrangian parcel under time-constant kinematics; returns
    (nlm (Nt+1, n) complex64, F (Nt+1, 3, 3), time (Nt+1), ugrad (3, 3))."""
    if "tau" not in DK:
        DK["tau"] = 1
    kind = DK["type"]
    if kind in ("simpleshear", "ss"):
        if dt is None:
            dt = sf.simpleshear_gamma_to_t(strain_target, DK["tau"]) / Nt
        F = np.array([sf.simpleshear_F(DK["plane"], DK["tau"], dt * nt) for nt in np.arange(Nt + 1)])
        ugrad = sf.simpleshear_ugrad(DK["plane"], DK["tau"])
    elif kind in ("pureshear", "ps"):
        if dt is None:
            dt = sf.pureshear_strainii_to_t(strain_target, DK["tau"]) / Nt
        F = np.array([sf.pureshear_F(DK["axis"], DK["q"], DK["tau"], dt * nt) for nt in np.arange(Nt + 1)])
        ugrad = sf.pureshear_ugrad(DK["axis"], DK["q"], DK["tau"])
    elif kind in ("rigidrotation", "rr"):
        raise Exception('type="rigidrotation" not yet supported')
    else:
        raise ValueError('Mode of deformation type="%s" not supported' % kind)

    time = dt * np.arange(Nt + 1)
    n = sf.nlm_len()
    L = sf.Lcap
    nlm = np.zeros((Nt + 1, n), dtype=np.complex64)
    if nlm0 is None:
        nlm[0, 0] = 1 / np.sqrt(4 * np.pi)
    else:
        nlm[0, :] = nlm0

    procs = []
    if iota is not None:
        procs.append("LROT")
    reg = None
    if nu is not None:
        procs.append("REG")
        reg = _reg_law(sf, L, nu, regexpo)
    if Lambda is not None:
        procs.append("CDRX")
    Gamma0_ = Gamma0(ugrad=ugrad) if callable(Gamma0) else Gamma0
    if Gamma0_ is not None:
        procs.append("DDRX")

    fb = FabricBatch(L, nlm=nlm[0:1].astype(np.complex128), ugrad=np.asarray(ugrad, dtype=np.float64)[None],
                     gam0=[float(Gamma0_) if Gamma0_ is not None else 0.0],
                     lam=[float(Lambda) if Lambda is not None else 0.0])
    for nt in range(1, Nt + 1):
        if procs:
            fb.evolve(1, dt, scheme="euler", procs=tuple(procs), iota=float(iota if iota is not None else 0.0),
                      zeta=float(zeta), reg=reg)
        nlm[nt, :] = fb.nlm()[0].cpu().numpy()
        # the trajectory is stored in complex64: the next step starts from it
        fb.set_nlm(torch.from_numpy(nlm[nt:nt + 1].astype(np.complex128)))
    torch.cuda.synchronize()
    return nlm, F, time, ugrad
