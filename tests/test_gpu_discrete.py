"""GPU parity of the discrete-grain path against the reference's outputs
(tests/golden/discrete.npz, written by oracle/_ref) and the numpy
restatement (oracle/discrete.py):

  * ri_LROT (src/dynamics.f90:112-137): the specfabpy face and the batched
    device face equal the reference trajectories to 1e-13 (40 Euler steps,
    iota = 1, -1, 0.3); 2^20 grains over 20 steps stay unit vectors and
    equal the restatement on a sample;
  * ri_to_nlm (src/moments.f90:98-131): L = 2, 4 (a2 / a4 branches), 6, 8,
    10 (a6 branch, rows above l = 6 zero) equal the reference to 1e-13
    relative to the largest coefficient; single grains equal conj(Y_lm);
    FabricBatch.set_grains on 4096 x 1000, 8192 x 100 and 2048 x 1024
    ensembles x grains (32, 4 and 64 lanes per ensemble), 1024 x 1024 and
    16 x 20000 (a workgroup per ensemble) equals the restatement (a6 branch)
    and the
    library's own a2_to_nlm /
    a4_to_nlm of the restated moments (a2 / a4 branches); odd N, partial
    waves and an unaligned output take the per-lane store path of ri_LROT
    and equal the restatement; an empty ensemble gives the
    reference's output for zero moments; L < 2 and odd L are rejected.
"""
import numpy as np
import pytest

from conftest import load_golden, relerr
from oracle import discrete as O

pytestmark = pytest.mark.gpu


def test_ri_LROT_matches_reference(gpu_ctx):
    import torch
    from specfab_amd import specfabpy as sf
    from specfab_amd.batch import ri_LROT_batch
    g = load_golden("discrete.npz")
    Nt = g["lrot_ri"].shape[1]
    for j, io in enumerate(g["lrot_iota"]):
        ref = g["lrot_ri"][j]
        got = sf.ri_LROT(g["lrot_ri0"], float(g["lrot_dt"]), Nt, g["lrot_D"], g["lrot_W"], io)
        assert np.abs(got - ref).max() < 1e-13, j
        dev = ri_LROT_batch(torch.from_numpy(g["lrot_ri0"]).cuda(), float(g["lrot_dt"]), Nt,
                            torch.from_numpy(g["lrot_D"]).cuda(), torch.from_numpy(g["lrot_W"]).cuda(), io)
        assert np.abs(dev.cpu().numpy() - ref).max() < 1e-13, j
    # Nt = 1: the trajectory is the initial state
    one = sf.ri_LROT(g["lrot_ri0"], 0.1, 1, g["lrot_D"][:1], g["lrot_W"][:1], 1.0)
    assert np.array_equal(one[0], g["lrot_ri0"])


def test_ri_LROT_many_grains(gpu_ctx):
    import torch
    from specfab_amd.batch import ri_LROT_batch
    from specfab_amd import workloads as W
    rng = np.random.default_rng(21)
    N, Nt, dt = 1 << 20, 21, 0.05
    r0 = rng.normal(size=(N, 3))
    r0 /= np.linalg.norm(r0, axis=1, keepdims=True)
    G = W.traceless_gradients(rng, Nt, sigma=1.0)
    D = (G + np.transpose(G, (0, 2, 1))) / 2
    Wm = (G - np.transpose(G, (0, 2, 1))) / 2
    ri = ri_LROT_batch(torch.from_numpy(r0).cuda(), dt, Nt, torch.from_numpy(D).cuda(),
                       torch.from_numpy(Wm).cuda(), 1.0)
    torch.cuda.synchronize()
    norms = torch.linalg.vector_norm(ri, dim=-1)
    assert float((norms - 1).abs().max()) < 1e-14
    sample = rng.choice(N, 4096, replace=False)
    ref = O.ri_LROT(r0[sample], dt, Nt, D, Wm, 1.0)
    assert np.abs(ri[:, torch.from_numpy(sample).cuda()].cpu().numpy() - ref).max() < 1e-13


def test_ri_LROT_ragged(gpu_ctx):
    """N odd (every other trajectory row starts 8 bytes off a 16-byte
    boundary), a partial last wave, and an output slice that starts 8 bytes
    off: the per-lane store path, against the restatement."""
    import torch
    from specfab_amd import specfabpy as sf
    rng = np.random.default_rng(23)
    g = load_golden("discrete.npz")
    for N in (1, 63, 65, 1001):
        r0 = rng.normal(size=(N, 3))
        r0 /= np.linalg.norm(r0, axis=1, keepdims=True)
        ref = O.ri_LROT(r0, 0.05, 9, g["lrot_D"][:9], g["lrot_W"][:9], 1.0)
        got = sf.ri_LROT(r0, 0.05, 9, g["lrot_D"][:9], g["lrot_W"][:9], 1.0)
        assert np.abs(got - ref).max() < 1e-14, N
    # device output 8 bytes past a 16-byte boundary (through the C-ABI)
    from specfab_amd._lib import check, lib
    from specfab_amd.batch import _ptr, _stream
    N, Nt = 256, 5
    r0 = rng.normal(size=(N, 3))
    r0 /= np.linalg.norm(r0, axis=1, keepdims=True)
    buf = torch.zeros(Nt * N * 3 + 1, dtype=torch.float64, device="cuda")
    D = torch.from_numpy(np.ascontiguousarray(g["lrot_D"][:Nt])).cuda()
    Wt = torch.from_numpy(np.ascontiguousarray(g["lrot_W"][:Nt])).cuda()
    r0d = torch.from_numpy(r0).cuda()
    check(lib().sfg_ri_LROT_batch(gpu_ctx.handle, N, Nt, _ptr(r0d), 0.05, _ptr(D), _ptr(Wt), 1.0, _ptr(buf[1:]),
                                  _stream()))
    ref = O.ri_LROT(r0, 0.05, Nt, g["lrot_D"][:Nt], g["lrot_W"][:Nt], 1.0)
    assert float(buf[0]) == 0.0
    assert np.abs(buf[1:].cpu().numpy().reshape(Nt, N, 3) - ref).max() < 1e-14


def test_ri_to_nlm_matches_reference(gpu_ctx):
    from specfab_amd import specfabpy as sf
    g = load_golden("discrete.npz")
    for L in g["r2n_L"]:
        for e in range(len(g["r2n_ri"])):
            ref = g["r2n_nlm%d" % L][e]
            got = sf.ri_to_nlm(g["r2n_ri"][e], g["r2n_wi"][e], int(L))
            assert got.shape == ref.shape
            assert relerr(got, ref) < 1e-13, (L, e)
    for r, ref in zip(g["r2n_one_ri"], g["r2n_one_nlm"]):
        assert np.abs(sf.ri_to_nlm(r[None, :], np.ones(1), 6) - ref).max() < 1e-14


def test_set_grains_batched(gpu_ctx):
    import torch
    from specfab_amd import specfabpy as sf
    from specfab_amd.batch import FabricBatch
    rng = np.random.default_rng(22)
    E, N = 4096, 1000
    ri = rng.normal(size=(E, N, 3)) * np.array([1.0, 0.6, 0.3])
    ri /= np.linalg.norm(ri, axis=-1, keepdims=True)
    wi = rng.uniform(0.5, 1.0, N)
    wi /= wi.sum()
    Rd = torch.from_numpy(ri).cuda()
    sample = [0, 1, 777, E - 1]
    # lanes per ensemble: 32 (4096 x 1000), 4 (8192 x 100), 64 (2048 x 1024);
    # a workgroup per ensemble (1024 x 1024, 16 x 20000)
    cases = ([(L, E, N) for L in (2, 4, 6, 12)] + [(L, 8192, 100) for L in (2, 4, 6, 8)] + [(6, 2048, 1024)]
             + [(6, 1024, 1024)] + [(L, 16, 20000) for L in (2, 4, 6)])
    for L, E_, N_ in cases:
        if (E_, N_) != (E, N):
            ri = rng.normal(size=(E_, N_, 3))
            ri /= np.linalg.norm(ri, axis=-1, keepdims=True)
            wi = rng.uniform(0.5, 1.0, N_)
            wi /= wi.sum()
            Rd = torch.from_numpy(ri).cuda()
            E, N = E_, N_
            sample = sorted({0, 1, min(777, E - 2), E - 1})
        b = FabricBatch(L, P=E)
        b.set_grains(Rd, wi)
        got = b.nlm().cpu().numpy()
        for e in sample:
            if L >= 6:
                ref = O.ri_to_nlm(ri[e], wi, L)
            else:
                ref = np.zeros_like(got[e])
                if L == 2:
                    ref[:6] = sf.a2_to_nlm(O.ri_moment(ri[e], wi, 2))
                else:
                    ref[:15] = sf.a4_to_nlm(O.ri_moment(ri[e], wi, 4))
            assert relerr(got[e], ref) < 1e-13, (L, e)


def test_ri_to_nlm_edges(gpu_ctx):
    from specfab_amd import specfabpy as sf
    from specfab_amd._lib import SFGError
    empty = np.zeros((0, 3))
    z6 = sf.ri_to_nlm(empty, np.zeros(0), 6)
    assert z6[0] == pytest.approx(1 / (2 * np.sqrt(np.pi)), rel=1e-15)
    assert np.all(z6[1:] == 0)
    z2 = sf.ri_to_nlm(empty, np.zeros(0), 2)
    assert relerr(z2, sf.a2_to_nlm(np.zeros((3, 3)))) < 1e-15
    for L in (0, 3):
        with pytest.raises(SFGError):
            sf.ri_to_nlm(np.array([[0.0, 0.0, 1.0]]), np.ones(1), L)
