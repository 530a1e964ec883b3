"""Pins for the reference fixture tests/golden/faces_L8.npz (written from the
reference's own Fortran by tests/golden/make_golden.py) behind the
nlm_LROT, eig3 / eigframe and Qnorm_tranisotropic faces:

  * nlm_LROT: row 0 is nlm0, and every row is the Euler step of the one
    before with the C oracle's M_LROT(D_j, W_j, iota, 0) (1e-13);
  * eig3: the eigenvalues are numpy's eigvalsh, largest first (1e-14), the
    vectors unit eigenvectors (M e = lam e to 1e-14), a right- or
    left-handed orthonormal set;
  * eigframe: 'ij' is eig3; 'xy' / 'xz' put last the eigenvector with the
    largest |z| / |y| component, keeping the other two in eig3's order;
  * Qnorm_tranisotropic: the C oracle's acoustic tensor (Reuss, Hill, Voigt)
    to 1e-12, and symmetric.
With the reference library present (oracle/_ref), the fixture is also
regenerated and compared bit for bit."""
import os

import numpy as np
import pytest

from conftest import ROOT, load_golden, relerr


@pytest.fixture(scope="module")
def orc():
    from oracle.oracle import Oracle
    return Oracle()


def test_nlm_LROT_is_euler_of_M_LROT(orc):
    g = load_golden("faces_L8.npz")
    X, D, W, dt = g["lrot_nlm"], g["lrot_D"], g["lrot_W"], float(g["lrot_dt"])
    assert np.array_equal(X[0], g["lrot_nlm0"])
    for j in range(len(X) - 1):
        M = orc.M_LROT(8, D[j], W[j], float(g["lrot_iota"]), 0.0)
        assert np.abs(X[j + 1] - (X[j] + dt * M @ X[j])).max() < 1e-13 * np.abs(X[j]).max()
    assert np.abs(X[-1] - X[0]).max() > 1e-3          # the fabric moved


def test_eig3_and_eigframe():
    g = load_golden("faces_L8.npz")
    for M, E, lam in zip(g["eig_M"], g["eig3_ei"], g["eig3_lami"]):
        assert np.abs(lam - np.linalg.eigvalsh(M)[::-1]).max() < 1e-14 * max(1, np.abs(lam).max())
        assert np.abs(M @ E.T - E.T * lam).max() < 1e-14 * max(1, np.abs(lam).max())
        assert np.abs(E @ E.T - np.eye(3)).max() < 1e-14
    for plane, k in (("xy", 2), ("xz", 1)):
        for E, lam, Ef, lf in zip(g["eig3_ei"], g["eig3_lami"], g["eigframe_%s_ei" % plane],
                                  g["eigframe_%s_lami" % plane]):
            i = int(np.argmax(np.abs(E[:, k])))
            order = [j for j in range(3) if j != i] + [i]
            assert np.array_equal(Ef, E[order]) and np.array_equal(lf, lam[order])
    assert np.array_equal(g["eigframe_ij_ei"], g["eig3_ei"])


def test_qnorm_matches_oracle(orc):
    g = load_golden("faces_L8.npz")
    for tag, alpha in (("reuss", 0.0), ("hill", 0.5), ("voigt", 1.0)):
        for x, Q in zip(g["qn_nlm"], g["qn_" + tag]):
            assert relerr(orc.Qnorm_tranisotropic(x, alpha, g["qn_Lame"]), Q) < 1e-12, tag
            assert np.abs(Q - Q.T).max() < 1e-12 * np.abs(Q).max()


def test_fixture_regenerates_bit_for_bit():
    from oracle import ref
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(ROOT, "tests", "golden",
                                                                              "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    new = mg.faces_fixture(ref.RefLib())
    old = load_golden("faces_L8.npz")
    assert set(new) == set(old)
    for k in new:
        assert np.array_equal(np.asarray(new[k]), old[k]), k


def test_evw_pins(orc):
    """Evw of the fixture: orthotropic grains against the restatement
    (oracle/orthotropic.py's Sachs tensor contracted with the state: 2e-6,
    the REAL(4) joint bodies), n' = 3 Sachs (alpha = 0) against
    oracle/nonlinear.py (1e-5: the REAL(4) a6 / a8 bodies' 1e-7, amplified
    where a random state's contraction nearly cancels; measured 1.8e-6)"""
    from oracle import nonlinear as N
    from oracle import orthotropic as O
    g = load_golden("faces_L8.npz")
    V, Wv, T = g["evw_v"], g["evw_w"], g["evw_tau"]
    iso = np.zeros(15, complex)
    iso[0] = 1 / np.sqrt(4 * np.pi)
    d42 = lambda A, B: np.einsum("lkij,ji->lk", A, B)
    d22 = lambda A, B: np.einsum("ij,ji->", A, B)
    Qi = O.sachs_Q(*O.ai(iso, iso, iso, orc)[1:], g["evwo_Eg"])
    for tag, use3 in (("given", True), ("derived", False)):
        for k, q in enumerate(g["evwo_q"]):
            Q = O.sachs_Q(*O.ai(q[0], q[1], q[2] if use3 else None, orc)[1:], g["evwo_Eg"])
            for s in range(len(V)):
                vw = np.outer(V[s], Wv[s])
                ref = d22(d42(Q, T[s]), vw) / d22(d42(Qi, T[s]), vw)
                assert abs(g["evwo_" + tag][k, s] - ref) < 2e-6 * max(1, abs(ref)), (tag, k, s)
    Si = N._structure(N.isotropic(8), orc)
    for k, x in enumerate(g["evw_nlm"]):
        S = N._structure(x, orc)
        for s in range(len(V)):
            vw = np.outer(V[s], Wv[s])
            ref = N._d22(N.sachs3(T[s], *S, (1, 1e3)), vw) / N._d22(N.sachs3(T[s], *Si, (1, 1e3)), vw)
            assert abs(g["evw_3_0_1"][k, s] / ref - 1) < 1e-5, (k, s)
