"""Pins for the orthotropic-grain oracle (oracle/orthotropic.py) and the
reference fixture tests/golden/ortho_L4.npz (written from the reference's
own Fortran by tests/golden/make_golden.py):

  * Eij_orthotropic (Sachs; v given and v = b x n derived; Cartesian and a
    random frame; two grain parameter sets) and Vi_elastic_orthotropic
    (Voigt, Abramson (1997) olivine, six directions; the reference's
    Cardano roots in its (qS1, qS2, qP) order) of the restatement equal the
    reference to 2e-6 / 1e-7 relative to the largest entry: the REAL(4)
    literal level of the reference's joint-distribution bodies;
  * the Cardano roots are the eigenvalues of Q / rho to 1e-7 (its REAL(4)
    exponents), qP the largest;
  * the joint tensors obey their definitions: a4_joint contracted over its
    second pair is a2b (x) (I - a2n) ... in trace form <b b> (1 - <(b.n)^2>)
    normalised, a2_orth and a4_orth have unit trace, a4_jointcross traced on
    its first pair is a2_orth; an isotropic fabric has Eij = 1 and
    direction-independent Vi;
  * Cij_to_Lame_orthotropic inverts to the Cij of the Voigt tensor of a
    single crystal along the axes (C11 = lam11 + 2 mu1, C44 = (mu2 + mu3)/2 ...);
  * the discrete-grain forms (tests/golden/ortho_discrete.npz): Eij 1e-7, Vi
    1e-12; repeating every grain leaves Eij unchanged; grain moments of a
    random planar rotation ensemble approach their closed forms.
With the reference library present (oracle/_ref), the fixture is also
regenerated and compared bit for bit."""
import os

import numpy as np
import pytest

from conftest import ROOT, load_golden, relerr
from oracle import orthotropic as O


@pytest.fixture(scope="module")
def orc():
    from oracle.oracle import Oracle
    return Oracle()


def test_eij_orthotropic_matches_reference(orc):
    g = load_golden("ortho_L4.npz")
    I, R = np.eye(3), g["frame"]
    for gi, Eg in enumerate(g["Eg"]):
        for tag, use3 in (("given", True), ("derived", False)):
            for k, q in enumerate(g["q"]):
                q3 = q[2] if use3 else 0 * q[2]
                got = O.Eij(q[0], q[1], q3, I[0], I[1], I[2], Eg, orc)
                assert relerr(got, g["Eij_xyz_%s_%d" % (tag, gi)][k]) < 2e-6, (gi, tag, k)
                got = O.Eij(q[0], q[1], q3, R[0], R[1], R[2], Eg, orc)
                assert relerr(got, g["Eij_rot_%s_%d" % (tag, gi)][k]) < 2e-6, (gi, tag, k)
    for k, q in enumerate(g["q8"]):
        got = O.Eij(q[0], q[1], 0 * q[2], I[0], I[1], I[2], g["Eg"][0], orc)
        assert relerr(got, g["Eij_xyz_derived8"][k]) < 2e-6


def test_vi_orthotropic_matches_reference(orc):
    g = load_golden("ortho_L4.npz")
    for tag, use3 in (("given", True), ("derived", False)):
        for k, q in enumerate(g["q"]):
            got = O.Vi(q[0], q[1], q[2] if use3 else None, g["Lame"], float(g["rho"]), g["theta"], g["phi"], orc)
            ref = g["Vi_" + tag][k].T
            assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-7, (tag, k)
    q = g["q8"][0]
    got = O.Vi(q[0], q[1], q[2], g["Lame"], float(g["rho"]), g["theta"], g["phi"], orc)
    assert np.abs(got - g["Vi_given8"][0].T).max() / got.max() < 1e-7


def test_cardano_against_eigenvalues():
    """the reference's Cardano roots are the eigenvalues of Q / rho (to the
    1e-8 its REAL(4) exponents leave), qP the largest"""
    rng = np.random.default_rng(12)
    for _ in range(20):
        G = rng.normal(size=(3, 3))
        Q = 1e11 * (np.eye(3) + 0.3 * (G + G.T))
        if np.linalg.eigvalsh(Q).min() <= 0:
            continue
        v = O.qnorm_to_vi(Q, 3000.0)
        ev = np.sqrt(np.linalg.eigvalsh(Q) / 3000.0)
        assert np.abs(np.sort(v) - ev).max() / ev.max() < 1e-7
        assert v[2] == v.max()
    # exact float exponents would reproduce eigvalsh to rounding: the REAL(4) ones are visible
    Q = np.diag([1.0, 2.0, 4.0]) * 1e11
    v = O.qnorm_to_vi(Q, 1.0)
    assert np.abs(np.sort(v) - np.sqrt([1e11, 2e11, 4e11])).max() / v.max() > 1e-12


def test_joint_tensor_identities(orc):
    from specfab_amd import workloads as W
    rng = np.random.default_rng(4)
    from oracle import nonlinear as N
    # exact moments (quadrature), so that the trace identities hold to rounding
    b, n = [(m[2], m[4], m[4]) for m in (N.moments(W.perturbed_isotropic(rng, 1, 4, sigma=0.15)[0], (2, 4))
                                         for _ in range(2))]
    J = O.a4_joint(b, n)
    # <b b n n>: tracing the n pair gives <b b> (the kernel's weight normalised away)
    s = 1 - np.einsum("ij,ij", b[0], n[0])
    assert np.abs(np.einsum("ijkk->ij", J) - (b[0] - np.einsum("ijpq,pq->ij", b[2], n[0])) / s).max() < 1e-14
    assert abs(np.trace(O.a2_orth(b, n)) - 1) < 1e-14
    assert abs(np.einsum("iijj", O.a4_orth(b, n)) - 1) < 1e-13
    C = O.a4_jointcross(b, n)                          # <n n v v>: tracing n n gives <v v>
    assert np.abs(np.einsum("iikl->kl", C) - O.a2_orth(b, n)).max() < 1e-14


def test_isotropic_fabric(orc):
    g = load_golden("ortho_L4.npz")
    iso = np.zeros(15, complex)
    iso[0] = 1 / np.sqrt(4 * np.pi)
    R = g["frame"]
    for Eg in g["Eg"]:
        assert np.abs(O.Eij(iso, iso, iso, R[0], R[1], R[2], Eg, orc) - 1).max() < 1e-12
    V = O.Vi(iso, iso, None, g["Lame"], float(g["rho"]), g["theta"], g["phi"], orc)
    assert np.abs(V - V[0]).max() / V.max() < 1e-7      # REAL(4) Wigner normalisations


def test_cij_to_lame_orthotropic():
    g = load_golden("ortho_L4.npz")
    C, L = g["Cij"], g["Lame"]
    mu = L[6:]
    assert np.allclose(C[:3], L[:3] + 2 * mu, rtol=1e-15)                        # C11 = lam11 + 2 mu1 ...
    # C44 = (mu2 + mu3) / 2, C55 = (mu1 + mu3) / 2, C66 = (mu1 + mu2) / 2
    assert np.allclose(C[3:6], [(mu[1] + mu[2]) / 2, (mu[0] + mu[2]) / 2, (mu[0] + mu[1]) / 2], rtol=1e-15)
    assert np.array_equal(C[6:], L[3:6])


def test_discrete_matches_reference(orc):
    """Eij_orthotropic_discrete / Vi_elastic_orthotropic__discrete of four
    grain ensembles (64, 64, 100 and 1 grains): Eij to 1e-7 (the continuum
    isotropic normalisation carries the REAL(4) joint bodies: a common factor
    1 + O(1e-8)), Vi to 1e-12 (exact grain moments, rotate_vector, Cardano)"""
    g = load_golden("ortho_discrete.npz")
    I, R = np.eye(3), g["frame"]
    for k in range(4):
        T = g["grains_%d" % k]
        for gi, Eg in enumerate(g["Eg"]):
            for tag, e in (("xyz", I), ("rot", R)):
                got = O.Eij_discrete(T[:, 0], T[:, 1], T[:, 2], e[0], e[1], e[2], Eg, orc)
                assert relerr(got, g["Eij_%s_%d" % (tag, gi)][k]) < 1e-7, (k, gi, tag)
        V = O.Vi_discrete(T[:, 0], T[:, 1], T[:, 2], g["Lame"], float(g["rho"]), g["theta"], g["phi"])
        assert np.abs(V - g["Vi"][k].T).max() / V.max() < 1e-12, k
    # an ensemble of one grain repeated five times has that grain's Eij
    T = g["grains_3"]
    rep = np.repeat(T, 5, axis=0)
    for Eg in g["Eg"]:
        a = O.Eij_discrete(T[:, 0], T[:, 1], T[:, 2], I[0], I[1], I[2], Eg, orc)
        b = O.Eij_discrete(rep[:, 0], rep[:, 1], rep[:, 2], I[0], I[1], I[2], Eg, orc)
        assert np.abs(a - b).max() < 1e-13 * np.abs(a).max()


def test_discrete_grains_are_a_sample_of_the_distribution(orc):
    """Grain triads rotated from a fixed crystal by random rotations about z:
    the ensemble's moments converge to the continuum moments of its axes
    (b, n uniformly in the x-y plane, v along z), so Vi of many grains equals
    Vi of the corresponding transversely isotropic moments; checked through
    the grains' moments themselves against the closed forms <b_x^2> = 1/2,
    <b_x^4> = 3/8, <v_z^4> = 1"""
    rng = np.random.default_rng(5)
    N = 20000
    t = rng.uniform(0, 2 * np.pi, N)
    b = np.stack([np.cos(t), np.sin(t), 0 * t], 1)
    n = np.stack([-np.sin(t), np.cos(t), 0 * t], 1)
    v = np.cross(b, n)
    a2, a4, jk = O.ai_discrete(b, n, v)
    assert abs(a2[0][0, 0] - 0.5) < 2e-2 and abs(a4[0][0, 0, 0, 0] - 3 / 8) < 2e-2
    assert abs(a4[2][2, 2, 2, 2] - 1) < 1e-14 and abs(jk[1][0, 0, 2, 2] - 0.5) < 2e-2


def test_fixture_regenerates_bit_for_bit():
    from oracle import ref
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(ROOT, "tests", "golden",
                                                                              "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    for fn, name in ((mg.ortho_fixture, "ortho_L4.npz"), (mg.ortho_discrete_fixture, "ortho_discrete.npz")):
        new = fn(ref.RefLib())
        old = load_golden(name)
        assert set(new) == set(old)
        for k in new:
            assert np.array_equal(np.asarray(new[k]), old[k]), (name, k)
