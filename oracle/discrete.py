"""Numpy restatement of the discrete-grain path.

Test infrastructure: imported only by tests/ as the checker; nothing under
specfab_amd/ imports it.

  ri_LROT     src/dynamics.f90:112-137: Euler steps of each grain axis under
              r' = (W + iota (r r^T D - D r r^T)) r, renormalised every step.
  ri_moment   src/moments.f90:98-131: the weighted k-th moment tensor
              sum_g w_g r_g^(x k) (k = 2, 4, 6) that ri_to_nlm hands to
              a2_to_nlm / a4_to_nlm / a6_to_nlm.
  a6_to_nlm   src/moments.f90:86-92: derived here numerically, independently of
              the sympy derivation in specfab_amd/tables/make_a6_map.py: on the
              unit sphere the degree-6 monomials span the harmonics l <= 6
              (28 = 1 + 5 + 9 + 13 functions), so conj(Y_lm) has unique
              monomial coefficients, found by a least-squares fit of scipy's
              sph_harm_y at 400 points of the sphere.  The reference evaluates
              the rows whose normalising root is sqrt(q) with q a half-integer
              (Sqrt(7.5), Sqrt(2.5), Sqrt(17.5), Sqrt(136.5), Sqrt(45.5) in
              src/include/a6_to_nlm__body.f90) with that root as a default-kind
              REAL(4) literal, in the rows of 2.5, 17.5 and 45.5 multiplied by
              3 (3/4, 3/8) in REAL(4) before it meets the FP64 sum: HALF_ROOT
              scales those rows by float32(c float32(sqrt(q))) / (c sqrt(q)),
              c = 1 or 3.  n00 is 1 / (2 sqrt(pi)) whatever the weights (the
              body's first line).
  ri_to_nlm   the L >= 6 branch of src/moments.f90:98-131 (rows above l = 6
              zero); L = 2 and 4 go through the reference's a2_to_nlm /
              a4_to_nlm, which the tests take from the library under test
              (pinned separately) with ri_moment's tensor.
"""
import numpy as np
from scipy.special import sph_harm_y

# (l, |m|) rows evaluated with a REAL(4) root sqrt(q), q a half-integer
HALF_ROOT = {(2, 1): (7.5, 1), (2, 2): (7.5, 1), (4, 2): (2.5, 3), (4, 4): (17.5, 3), (6, 1): (136.5, 1),
             (6, 4): (45.5, 3)}


def ri_LROT(ri0, dt, Nt, D, W, iota):
    """ri (Nt, N, 3) from ri0 (N, 3), D and W (Nt, 3, 3) (src/dynamics.f90:112-137)."""
    r = np.array(ri0, dtype=np.float64)
    out = np.zeros((Nt,) + r.shape)
    out[0] = r
    for t in range(Nt - 1):
        mm = r[:, :, None] * r[:, None, :]
        Wp = iota * (mm @ D[t] - D[t] @ mm)
        dri = np.einsum("gij,gj->gi", W[t] + Wp, r)
        rn = r + dt * dri
        r = rn / np.linalg.norm(rn, axis=1, keepdims=True)
        out[t + 1] = r
    return out


def ri_moment(ri, wi, k):
    """sum_g w_g r_g^(x k) as a (3,)*k array (src/moments.f90:106-127)."""
    r = np.asarray(ri, dtype=np.float64)
    w = np.asarray(wi, dtype=np.float64)
    letters = "abcdef"[:k]
    spec = ",".join("g" + c for c in letters)
    return np.einsum("g," + spec + "->" + letters, w, *([r] * k))


def lm6():
    return [(l, m) for l in range(0, 7, 2) for m in range(-l, l + 1)]


def _a6_matrix():
    """D (28 nlm rows, 729 a6 entries): n_lm = D @ a6.ravel() for a6 of unit axes."""
    rng = np.random.default_rng(6)
    v = rng.normal(size=(400, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    mons = [(a, b, 6 - a - b) for a in range(6, -1, -1) for b in range(6 - a, -1, -1)]
    A = np.stack([v[:, 0] ** a * v[:, 1] ** b * v[:, 2] ** c for a, b, c in mons], axis=1)
    th = np.arccos(np.clip(v[:, 2], -1, 1))
    ph = np.arctan2(v[:, 1], v[:, 0])
    Y = np.stack([np.conj(sph_harm_y(l, m, th, ph)) for l, m in lm6()], axis=1)
    C, *_ = np.linalg.lstsq(A, Y, rcond=None)          # (28 monomials, 28 rows)
    # spread each monomial's coefficient evenly over the a6 entries that hold it
    D = np.zeros((28, 729), dtype=np.complex128)
    idx = np.indices((3,) * 6).reshape(6, -1).T
    cnt = np.stack([(idx == 0).sum(1), (idx == 1).sum(1), (idx == 2).sum(1)], axis=1)
    mult = {}
    for e, (a, b, c) in enumerate(cnt):
        mult[(a, b, c)] = mult.get((a, b, c), 0) + 1
    for e, (a, b, c) in enumerate(cnt):
        D[:, e] = C[mons.index((a, b, c))] / mult[(a, b, c)]
    for row, (l, m) in enumerate(lm6()):
        if (l, abs(m)) in HALF_ROOT:
            q, c = HALF_ROOT[(l, abs(m))]
            D[row] *= float(np.float32(c) * np.float32(np.sqrt(q))) / (c * np.sqrt(q))
    return D


_D6 = None


def a6_to_nlm(a6):
    """nlm (28,) of the l <= 6 coefficients from a6 (src/moments.f90:86-92)."""
    global _D6
    if _D6 is None:
        _D6 = _a6_matrix()
    z = _D6 @ np.asarray(a6, dtype=np.float64).ravel()
    z[0] = 1 / (2 * np.sqrt(np.pi))
    return z


def ri_to_nlm(ri, wi, L):
    """L >= 6 branch of ri_to_nlm (src/moments.f90:98-131)."""
    if L < 6:
        raise ValueError("the L = 2, 4 branches take the reference's a2_to_nlm / a4_to_nlm")
    out = np.zeros((L + 1) * (L + 2) // 2, dtype=np.complex128)
    out[:28] = a6_to_nlm(ri_moment(ri, wi, 6))
    return out
