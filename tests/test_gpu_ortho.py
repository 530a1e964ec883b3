"""GPU parity of the orthotropic-grain (olivine) homogenization against the
reference's outputs (tests/golden/ortho_L4.npz, written by oracle/_ref) and
the numpy restatement (oracle/orthotropic.py):

  * Eij_orthotropic (Sachs): v given and v = b x n derived, Cartesian and a
    random frame, two grain parameter sets, single-parcel and batched faces:
    2e-6 relative to the largest entry against the reference (the REAL(4)
    literal level of its joint-distribution bodies), 1e-11 against the
    restatement, which shares the body a4;
  * Vi_elastic_orthotropic (Voigt, Abramson (1997) olivine, six directions):
    1e-7 relative against the reference in its own (qS1, qS2, qP) order,
    1e-11 against the restatement (the same Cardano solution), single and
    batched;
  * L = 8 states (only the l <= 4 blocks enter) and a batch of 4096 random
    triples against the restatement on a sample;
  * Cij_to_Lame_orthotropic, and the errors for mismatched batches;
  * discrete grains (tests/golden/ortho_discrete.npz: ensembles of 64, 64,
    100 and 1 grains): Eij 1e-7 against the reference (the continuum
    isotropic normalisation) and 1e-11 against the restatement, Vi 1e-11
    against both (the acoustic tensor formed in the grains' frame for k =
    the direction, the reference's rotated grains for k = z: the same
    invariants); single-ensemble and batched faces, a 4096-grain ensemble
    split over several workgroups.
"""
import numpy as np
import pytest

from conftest import load_golden, relerr
from oracle import orthotropic as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def orc():
    from oracle.oracle import Oracle
    return Oracle()


def test_eij_orthotropic_single(gpu_ctx, orc):
    from specfab_amd import specfabpy as sf
    g = load_golden("ortho_L4.npz")
    I, R = np.eye(3), g["frame"]
    sf.init(4)
    for gi, Eg in enumerate(g["Eg"]):
        for tag, use3 in (("given", True), ("derived", False)):
            for k, q in enumerate(g["q"]):
                q3 = q[2] if use3 else 0 * q[2]
                for ftag, fr in (("xyz", I), ("rot", R)):
                    got = sf.Eij_orthotropic(q[0], q[1], q3, fr[0], fr[1], fr[2], Eg, 0.0, 1)
                    assert relerr(got, g["Eij_%s_%s_%d" % (ftag, tag, gi)][k]) < 2e-6, (gi, tag, k, ftag)
                    ora = O.Eij(q[0], q[1], q3, fr[0], fr[1], fr[2], Eg, orc)
                    assert relerr(got, ora) < 1e-11, (gi, tag, k, ftag)
    sf.init(8)
    for k, q in enumerate(g["q8"]):
        got = sf.Eij_orthotropic(q[0], q[1], 0 * q[2], I[0], I[1], I[2], g["Eg"][0], 0.0, 1)
        assert relerr(got, g["Eij_xyz_derived8"][k]) < 2e-6


def test_eij_orthotropic_batched(gpu_ctx, orc):
    import torch
    from specfab_amd import specfabpy as sf
    from specfab_amd import workloads as W
    from specfab_amd.batch import FabricBatch, eij_orthotropic
    g = load_golden("ortho_L4.npz")
    Q, R = g["q"], g["frame"]
    K = len(Q)
    for gi, Eg in enumerate(g["Eg"]):
        for tag, use3 in (("given", True), ("derived", False)):
            q3 = Q[:, 2] if use3 else None
            Ex = sf.Eij_orthotropic_arr(Q[:, 0], Q[:, 1], q3, np.tile(np.eye(3)[0], (K, 1)),
                                        np.tile(np.eye(3)[1], (K, 1)), np.tile(np.eye(3)[2], (K, 1)), Eg, 0.0, 1)
            assert relerr(Ex, g["Eij_xyz_%s_%d" % (tag, gi)]) < 2e-6, (gi, tag)
            Er = sf.Eij_orthotropic_arr(Q[:, 0], Q[:, 1], q3, np.tile(R[0], (K, 1)), np.tile(R[1], (K, 1)),
                                        np.tile(R[2], (K, 1)), Eg, 0.0, 1)
            assert relerr(Er, g["Eij_rot_%s_%d" % (tag, gi)]) < 2e-6, (gi, tag)
    # zero v states derive v = b x n like a missing third batch
    b1, b2 = FabricBatch(4, nlm=Q[:, 0]), FabricBatch(4, nlm=Q[:, 1])
    b0 = FabricBatch(4, nlm=0 * Q[:, 2])
    E0 = eij_orthotropic(b1, b2, b0, g["Eg"][1])
    En = eij_orthotropic(b1, b2, None, g["Eg"][1])
    torch.cuda.synchronize()
    assert relerr(E0.cpu().numpy(), g["Eij_xyz_derived_1"]) < 2e-6
    assert torch.equal(E0, En)
    # a larger batch of random triples against the restatement on a sample
    rng = np.random.default_rng(83)
    P = 4096
    X = [W.perturbed_isotropic(rng, P, 4, sigma=0.15) for _ in range(3)]
    bs = [FabricBatch(4, nlm=x) for x in X]
    I = np.eye(3)
    for use3 in (True, False):
        E = eij_orthotropic(bs[0], bs[1], bs[2] if use3 else None, g["Eg"][1]).cpu().numpy()
        for k in (0, 1, 2047, P - 1):
            ora = O.Eij(X[0][k], X[1][k], X[2][k] if use3 else 0 * X[2][k], I[0], I[1], I[2], g["Eg"][1], orc)
            assert relerr(E[k], ora) < 1e-11, (use3, k)


def test_vi_orthotropic(gpu_ctx, orc):
    import torch
    from specfab_amd import specfabpy as sf
    from specfab_amd.batch import FabricBatch, vi_elastic_orthotropic
    g = load_golden("ortho_L4.npz")
    Lame, rho, th, ph = g["Lame"], float(g["rho"]), g["theta"], g["phi"]
    Q = g["q"]
    sf.init(4)
    for tag, use3 in (("given", True), ("derived", False)):
        for k, q in enumerate(Q):
            got = sf.Vi_elastic_orthotropic(q[0], q[1], q[2] if use3 else None, 1.0, Lame, rho, th, ph)
            ref = g["Vi_" + tag][k]
            assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-7, (tag, k)
            ora = O.Vi(q[0], q[1], q[2] if use3 else None, Lame, rho, th, ph, orc)
            assert np.abs(got.T - ora).max() / ora.max() < 1e-11, (tag, k)
        b = [FabricBatch(4, nlm=Q[:, i]) for i in range(3)]
        V = vi_elastic_orthotropic(b[0], b[1], b[2] if use3 else None, Lame, rho, th, ph)
        torch.cuda.synchronize()
        V = V.cpu().numpy()                                   # (K, ndir, 3)
        ref = np.transpose(g["Vi_" + tag], (0, 2, 1))
        assert np.abs(V - ref).max() / np.abs(ref).max() < 1e-7, tag
    sf.init(8)
    q = g["q8"][0]
    got = sf.Vi_elastic_orthotropic(q[0], q[1], q[2], 1.0, Lame, rho, th, ph)
    assert np.abs(got - g["Vi_given8"][0]).max() / np.abs(g["Vi_given8"][0]).max() < 1e-7


def test_cij_to_lame_orthotropic_and_errors(gpu_ctx):
    from specfab_amd import specfabpy as sf
    from specfab_amd.batch import FabricBatch, eij_orthotropic
    g = load_golden("ortho_L4.npz")
    assert relerr(sf.Cij_to_Lame_orthotropic(g["Cij"]), g["Lame"]) < 1e-15
    b4 = FabricBatch(4, nlm=g["q"][:, 0])
    b4s = FabricBatch(4, nlm=g["q"][:2, 1])
    with pytest.raises(ValueError):
        eij_orthotropic(b4, b4s, None, g["Eg"][0])
    with pytest.raises(ValueError):
        sf.Eij_orthotropic(g["q"][0, 0], g["q8"][0, 1], None, *np.eye(3), g["Eg"][0], 0.0, 1)


def test_discrete_grains(gpu_ctx, orc):
    import torch
    from specfab_amd import specfabpy as sf
    from specfab_amd.batch import eij_orthotropic_discrete, vi_elastic_orthotropic_discrete
    g = load_golden("ortho_discrete.npz")
    I, R = np.eye(3), g["frame"]
    Lame, rho, th, ph = g["Lame"], float(g["rho"]), g["theta"], g["phi"]
    for k in range(4):
        T = g["grains_%d" % k]
        b, n, v = T[:, 0], T[:, 1], T[:, 2]
        for gi, Eg in enumerate(g["Eg"]):
            for tag, e in (("xyz", I), ("rot", R)):
                got = sf.Eij_orthotropic_discrete(b, n, v, e[0], e[1], e[2], Eg, 0.0, 1)
                assert relerr(got, g["Eij_%s_%d" % (tag, gi)][k]) < 1e-7, (k, gi, tag)
                assert relerr(got, O.Eij_discrete(b, n, v, e[0], e[1], e[2], Eg, orc)) < 1e-11, (k, gi, tag)
        V = sf.Vi_elastic_orthotropic__discrete(np.transpose(T, (1, 2, 0)), 1.0, Lame, rho, th, ph)
        assert np.abs(V - g["Vi"][k]).max() / V.max() < 1e-11, k
    # batched: the two 64-grain ensembles in one launch
    T = np.stack([g["grains_0"], g["grains_1"]])                      # (2, 64, 3, 3)
    dev = torch.device("cuda:0")
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    E = eij_orthotropic_discrete(t(T[:, :, 0]), t(T[:, :, 1]), t(T[:, :, 2]), g["Eg"][1],
                                 frames=t(np.stack([R, R])))
    V = vi_elastic_orthotropic_discrete(t(T[:, :, 0]), t(T[:, :, 1]), t(T[:, :, 2]), Lame, rho, th, ph)
    torch.cuda.synchronize()
    assert relerr(E.cpu().numpy(), g["Eij_rot_1"][:2]) < 1e-7
    assert np.abs(V.cpu().numpy() - np.transpose(g["Vi"][:2], (0, 2, 1))).max() / V.max().item() < 1e-11
    Ea = sf.Eij_orthotropic_discrete_arr(T[:, :, 0], T[:, :, 1], T[:, :, 2], np.tile(I[0], (2, 1)),
                                         np.tile(I[1], (2, 1)), np.tile(I[2], (2, 1)), g["Eg"][0], 0.0, 1)
    assert relerr(Ea, g["Eij_xyz_0"][:2]) < 1e-7
    # a large ensemble: grain sums over several workgroups, against the restatement
    rng = np.random.default_rng(9)
    Nb = 4096
    Qs = np.linalg.qr(np.eye(3) + 0.5 * rng.normal(size=(Nb, 3, 3)))[0]
    Tb = np.transpose(Qs, (0, 2, 1))
    got = sf.Eij_orthotropic_discrete(Tb[:, 0], Tb[:, 1], Tb[:, 2], R[0], R[1], R[2], g["Eg"][1], 0.0, 1)
    ora = O.Eij_discrete(Tb[:, 0], Tb[:, 1], Tb[:, 2], R[0], R[1], R[2], g["Eg"][1], orc)
    assert relerr(got, ora) < 1e-11
    V = sf.Vi_elastic_orthotropic__discrete(np.transpose(Tb, (1, 2, 0)), 1.0, Lame, rho, th, ph)
    ora = O.Vi_discrete(Tb[:, 0], Tb[:, 1], Tb[:, 2], Lame, rho, th, ph)
    assert np.abs(V.T - ora).max() / ora.max() < 1e-11
    from specfab_amd._lib import SFGError
    with pytest.raises(SFGError):
        sf.Eij_orthotropic_discrete(Tb[:0, 0], Tb[:0, 1], Tb[:0, 2], *np.eye(3), g["Eg"][0], 0.0, 1)
