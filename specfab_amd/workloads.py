"""Seeded synthetic inputs for the benchmark configurations (numpy, host).

  lagrangian_parcel_c1   one parcel, L=8, uniaxial compression
                         ugrad = diag(0.5, 0.5, -1) * 1e-3, isotropic start
  eulerian_field         P parcels, ugrad_ij ~ N(0, 1e-3) iid made traceless,
                         Gamma0 ~ LogUniform(1e-4, 1e-2), isotropic start
                         perturbed by N(0, 1e-3) on every real l >= 2 DOF
                         (i.e. Re, and Im for m > 0, then n_l^-m =
                         (-1)^m conj(n_l^m) so the ODF stays real)
  uneven_batch           L ~ U{4, 8, 12, 20}, steps ~ U[50, 500],
                         Lambda ~ LogUniform(1e-4, 1e-3), else as above

States are returned as complex nlm (P, n) in specfab order.
"""
import numpy as np

ISO = 1 / np.sqrt(4 * np.pi)


def nlm_len(L):
    return (L + 1) * (L + 2) // 2


def lm_list(L):
    return [(l, m) for l in range(0, L + 1, 2) for m in range(-l, l + 1)]


def traceless_gradients(rng, P, sigma=1e-3):
    G = rng.normal(0.0, sigma, (P, 3, 3))
    tr = np.trace(G, axis1=1, axis2=2)
    G -= tr[:, None, None] / 3 * np.eye(3)
    return G


def perturbed_isotropic(rng, P, L, sigma=1e-3):
    lm = lm_list(L)
    n = len(lm)
    nlm = np.zeros((P, n), dtype=np.complex128)
    nlm[:, 0] = ISO
    idx = {v: i for i, v in enumerate(lm)}
    for i, (l, m) in enumerate(lm):
        if l < 2 or m < 0:
            continue
        re = rng.normal(0.0, sigma, P)
        im = rng.normal(0.0, sigma, P) if m > 0 else np.zeros(P)
        nlm[:, i] = re + 1j * im
        if m > 0:
            nlm[:, idx[(l, -m)]] = (-1) ** m * np.conj(nlm[:, i])
    return nlm


def lagrangian_parcel_c1():
    L = 8
    nlm = np.zeros((1, nlm_len(L)), dtype=np.complex128)
    nlm[0, 0] = ISO
    ugrad = (np.diag([0.5, 0.5, -1.0]) * 1e-3)[None]
    return dict(L=L, nlm=nlm, ugrad=ugrad, gam0=np.zeros(1), lam=np.zeros(1), dt=1.0, nsteps=1000,
                scheme="euler", procs=("LROT", "REG"))


def eulerian_field(P, L=12, seed=1, nsteps=200, dt=0.5):
    rng = np.random.default_rng(seed)
    ugrad = traceless_gradients(rng, P)
    gam0 = 10.0 ** rng.uniform(-4, -2, P)
    nlm = perturbed_isotropic(rng, P, L)
    return dict(L=L, nlm=nlm, ugrad=ugrad, gam0=gam0, lam=np.zeros(P), dt=dt, nsteps=nsteps, scheme="rk4",
                procs=("LROT", "REG", "DDRX"))


def uneven_batch(P, seed=4, Ls=(4, 8, 12, 20), smin=50, smax=500, dt=0.5):
    rng = np.random.default_rng(seed)
    L = rng.choice(np.asarray(Ls), P)
    steps = rng.integers(smin, smax + 1, P)
    ugrad = traceless_gradients(rng, P)
    gam0 = 10.0 ** rng.uniform(-4, -2, P)
    lam = 10.0 ** rng.uniform(-4, -3, P)
    return dict(L=L, steps=steps, ugrad=ugrad, gam0=gam0, lam=lam, dt=dt, scheme="rk4",
                procs=("LROT", "REG", "DDRX", "CDRX"), seed=seed)
