"""The reduced-form oracle (oracle/oracle.py: nlm_to_rnlm, rnlm_to_nlm,
reduce_M; src/reducedform.f90:76-183) pinned against the reference's own
outputs (tests/golden/reduced_L8.npz, written from oracle/_ref), against the
reference library where it is built, and against what the mathematics fixes:
the ODF rebuilt from rnlm_to_nlm is real on the sphere, and the reduced
operators reproduce M nlm at the reduced rows for any M (the check of
tests/reduced-form/reduced-form.f90:57-111)."""
import numpy as np
import pytest

from conftest import load_golden
from oracle import oracle as O


def _sweep(L):
    return np.random.default_rng(100 + L)


def test_lengths_and_order():
    for L in range(0, 31, 2):
        assert O.rnlm_len(L) == len(O.rlm_list(L)) == sum(l + 1 for l in range(0, L + 1, 2))
    assert O.rlm_list(4)[:6] == [(0, 0), (2, 0), (2, 1), (2, 2), (4, 0), (4, 1)]


def test_vs_golden():
    g = load_golden("reduced_L8.npz")
    L = int(g["L"])
    for k, M in enumerate(g["M"]):
        red = O.reduce_M(M, L)
        for i, name in enumerate(("Mrr", "Mri", "Mir", "Mii")):
            assert np.array_equal(red[i], g[name][k]), name
    assert np.array_equal(O.nlm_to_rnlm(g["nlm"], L), g["rnlm_of_nlm"])
    assert np.array_equal(O.rnlm_to_nlm(g["rnlm"], L), g["nlm_of_rnlm"])


def test_vs_reference_library():
    from oracle import ref
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    r = ref.RefLib()
    for L in (2, 12, 20):
        n = r.init(L)
        rng = _sweep(L)
        M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        for a, b in zip(O.reduce_M(M, L), r.reduce_M(M)):
            assert np.array_equal(a, b)
        v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        assert np.array_equal(O.nlm_to_rnlm(v, L), r.nlm_to_rnlm(v))
        w = rng.standard_normal(O.rnlm_len(L)) + 1j * rng.standard_normal(O.rnlm_len(L))
        assert np.array_equal(O.rnlm_to_nlm(w, L), r.rnlm_to_nlm(w))


@pytest.mark.parametrize("L", [4, 8])
def test_rebuilt_odf_is_real(L):
    """n_l^{-m} = (-1)^m conj(n_l^m) is what makes sum n_lm Y_lm real; scipy's
    Y_l^m carry the same Condon-Shortley phase as specfab's."""
    from scipy.special import sph_harm_y
    rng = _sweep(L)
    r = rng.standard_normal(O.rnlm_len(L)) + 1j * rng.standard_normal(O.rnlm_len(L))
    r[[k for k, (l, m) in enumerate(O.rlm_list(L)) if m == 0]] = rng.standard_normal(L // 2 + 1)
    nlm = O.rnlm_to_nlm(r, L)
    th, ph = np.meshgrid(np.linspace(0.1, 3.0, 7), np.linspace(0.0, 6.0, 9))
    f = sum(c * sph_harm_y(l, m, th, ph) for c, (l, m) in zip(nlm, O.lm_list(L)))
    assert np.max(np.abs(f.imag)) < 1e-12 * np.max(np.abs(f.real))
    # a wrong sign on the odd negative-m terms leaves an imaginary part
    bad = nlm.copy()
    for j, (l, m) in enumerate(O.lm_list(L)):
        if m < 0 and m % 2:
            bad[j] = -bad[j]
    fb = sum(c * sph_harm_y(l, m, th, ph) for c, (l, m) in zip(bad, O.lm_list(L)))
    assert np.max(np.abs(fb.imag)) > 1e-3


@pytest.mark.parametrize("L", [2, 6, 12])
def test_reduced_operator_reproduces_full_product(L):
    rng = _sweep(L)
    n = O.nlm_len(L)
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    r = rng.standard_normal(O.rnlm_len(L)) + 1j * rng.standard_normal(O.rnlm_len(L))
    full = O.nlm_to_rnlm(M @ O.rnlm_to_nlm(r, L), L)
    Mrr, Mri, Mir, Mii = O.reduce_M(M, L)
    red = Mrr @ r.real + Mri @ r.imag + 1j * (Mir @ r.real + Mii @ r.imag)
    assert np.allclose(red, full, rtol=0, atol=1e-12 * np.abs(full).max())
    assert np.array_equal(O.nlm_to_rnlm(O.rnlm_to_nlm(r, L), L), r)


def test_reduced_lrot_from_c_oracle():
    """reduce_M of the C oracle's M_LROT reproduces M_LROT nlm for a real ODF
    (tests/reduced-form/reduced-form.f90:57-111)."""
    o = O.Oracle()
    L = 8
    rng = _sweep(L)
    eps = rng.standard_normal((3, 3))
    eps = eps + eps.T
    omg = rng.standard_normal((3, 3))
    omg = omg - omg.T
    M = o.M_LROT(L, eps, omg, 1.0, 0.0)
    r = rng.standard_normal(O.rnlm_len(L)) + 1j * rng.standard_normal(O.rnlm_len(L))
    r[0] = 1 / np.sqrt(4 * np.pi)
    nlm = O.rnlm_to_nlm(r, L)
    Mrr, Mri, Mir, Mii = O.reduce_M(M, L)
    red = Mrr @ r.real + Mri @ r.imag + 1j * (Mir @ r.real + Mii @ r.imag)
    assert np.allclose(red, O.nlm_to_rnlm(M @ nlm, L), rtol=0, atol=1e-12)
