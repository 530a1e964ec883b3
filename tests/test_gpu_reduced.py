"""GPU parity of the reduced real form (src/reducedform.f90:76-183): the
single-node specfabpy face against the reference's outputs
(tests/golden/reduced_L8.npz), the batched C-ABI entry points against the
oracle (bit-exact: the kernels only gather, negate and add), the Eulerian
export chain M_LROT_batch -> reduce_M_batch, and round trips at field size."""
import numpy as np
import pytest

from conftest import load_golden, relerr
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NAMES = ("Mrr", "Mri", "Mir", "Mii")


def test_single_node_face_vs_reference(gpu_ctx):
    from specfab_amd import specfabpy as sf
    g = load_golden("reduced_L8.npz")
    L = int(g["L"])
    _, n = sf.init(L)
    rn = sf.get_rnlm_len()
    assert rn == O.rnlm_len(L)
    for k, M in enumerate(g["M"]):
        red = sf.reduce_M(M, rn)
        for i, name in enumerate(NAMES):
            assert np.array_equal(red[i], g[name][k]), (k, name)
    for k in range(len(g["nlm"])):
        assert np.array_equal(sf.nlm_to_rnlm(g["nlm"][k], rn), g["rnlm_of_nlm"][k])
        assert np.array_equal(sf.rnlm_to_nlm(g["rnlm"][k], n), g["nlm_of_rnlm"][k])


@pytest.mark.parametrize("L", [0, 2, 8, 12, 20, 30])
def test_batched_vs_oracle(gpu_ctx, L):
    import torch
    from specfab_amd import batch as B
    print("reduced batch L=%d" % L, flush=True)
    rng = np.random.default_rng(500 + L)
    n, rn = O.nlm_len(L), O.rnlm_len(L)
    P = 37
    M = rng.standard_normal((P, n, n)) + 1j * rng.standard_normal((P, n, n))
    red = B.reduce_M_batch(L, torch.from_numpy(M).cuda())
    want = O.reduce_M(M, L)
    for i, name in enumerate(NAMES):
        assert np.array_equal(red[i].cpu().numpy(), want[i]), name
    v = rng.standard_normal((P, n)) + 1j * rng.standard_normal((P, n))
    assert np.array_equal(B.nlm_to_rnlm_batch(L, torch.from_numpy(v).cuda()).cpu().numpy(), O.nlm_to_rnlm(v, L))
    w = rng.standard_normal((P, rn)) + 1j * rng.standard_normal((P, rn))
    assert np.array_equal(B.rnlm_to_nlm_batch(L, torch.from_numpy(w).cuda()).cpu().numpy(), O.rnlm_to_nlm(w, L))


def test_eulerian_export_chain(gpu_ctx, oracle):
    """per-node M_LROT and M_DDRX_src built and reduced on the device, the
    assembly of src/specfabpy/fenics/CPO.py:200-201 for a whole mesh"""
    import torch
    from specfab_amd import batch as B
    L, P = 12, 64
    rng = np.random.default_rng(7)
    eps = rng.standard_normal((P, 3, 3))
    eps = eps + eps.transpose(0, 2, 1)
    omg = rng.standard_normal((P, 3, 3))
    omg = omg - omg.transpose(0, 2, 1)
    M = B.M_LROT_batch(L, torch.from_numpy(eps).cuda(), torch.from_numpy(omg).cuda(), 1.0, 0.0)
    red = [t.cpu().numpy() for t in B.reduce_M_batch(L, M)]
    Md = B.M_DDRX_src_batch(L, torch.from_numpy(eps).cuda())
    redd = [t.cpu().numpy() for t in B.reduce_M_batch(L, Md)]
    for p in range(0, P, 9):
        want = O.reduce_M(oracle.M_LROT(L, eps[p], omg[p], 1.0, 0.0), L)
        wantd = O.reduce_M(oracle.M_DDRX_src(L, eps[p]), L)
        for i in range(4):
            assert relerr(red[i][p], want[i]) < 1e-13
            assert relerr(redd[i][p], wantd[i]) < 1e-13


@pytest.mark.parametrize("L", [4, 12, 20])
def test_fused_export_equals_chain(gpu_ctx, oracle, L):
    """M_LROT_reduced_batch / M_DDRX_src_reduced_batch (the reduced blocks as
    a linear map of each node's coupling weights, the full matrices never
    formed) equal the two-kernel chain to 1e-14 relative to the node's
    largest entry (the same sums in another order), and the oracle's
    reduce_M of its operators to 1e-13"""
    import torch
    from specfab_amd import batch as B
    print("fused export L=%d" % L, flush=True)
    P = 300
    rng = np.random.default_rng(70 + L)
    eps = rng.standard_normal((P, 3, 3))
    eps = eps + eps.transpose(0, 2, 1)
    omg = rng.standard_normal((P, 3, 3))
    omg = omg - omg.transpose(0, 2, 1)
    E, W = torch.from_numpy(eps).cuda(), torch.from_numpy(omg).cuda()
    def same(c, f, tag):
        c, f = c.cpu().numpy(), f.cpu().numpy()
        scale = np.abs(c).reshape(len(c), -1).max(axis=1)[:, None, None]
        assert (np.abs(f - c) <= 1e-14 * scale).all(), tag

    for zeta in (0.0, 0.4):
        chain = B.reduce_M_batch(L, B.M_LROT_batch(L, E, W, 1.0, zeta))
        fused = B.M_LROT_reduced_batch(L, E, W, 1.0, zeta)
        for i in range(4):
            same(chain[i], fused[i], (zeta, NAMES[i]))
    chain = B.reduce_M_batch(L, B.M_DDRX_src_batch(L, E))
    fused = B.M_DDRX_src_reduced_batch(L, E)
    for i in range(4):
        same(chain[i], fused[i], NAMES[i])
    fl = [t.cpu().numpy() for t in B.M_LROT_reduced_batch(L, E, W, 1.0, 0.0)]
    for p in (0, 151, P - 1):
        want = O.reduce_M(oracle.M_LROT(L, eps[p], omg[p], 1.0, 0.0), L)
        for i in range(4):
            assert relerr(fl[i][p], want[i]) < 1e-13


def test_fused_export_many_chunks(gpu_ctx):
    """more 16-node chunks than grid rows (16 x 65535 + 37 nodes at L = 2):
    workgroups loop over several chunks, the last one partial; equal to the
    chain to 1e-14 of each node's largest entry"""
    import torch
    from specfab_amd import batch as B
    print("fused export, 16 x 65535 + 37 nodes", flush=True)
    P = 16 * 65535 + 37
    g = torch.Generator(device="cuda").manual_seed(5)
    E = torch.randn((P, 3, 3), device="cuda", dtype=torch.float64, generator=g)
    E = E + E.transpose(1, 2)
    W = torch.randn((P, 3, 3), device="cuda", dtype=torch.float64, generator=g)
    W = W - W.transpose(1, 2)
    for tag, chain, fused in (("LROT", lambda: B.reduce_M_batch(2, B.M_LROT_batch(2, E, W, 1.0, 0.3)),
                               lambda: B.M_LROT_reduced_batch(2, E, W, 1.0, 0.3)),
                              ("DDRX", lambda: B.reduce_M_batch(2, B.M_DDRX_src_batch(2, E)),
                               lambda: B.M_DDRX_src_reduced_batch(2, E))):
        c, f = chain(), fused()
        for i in range(4):
            scale = c[i].abs().reshape(P, -1).amax(dim=1)[:, None, None]
            assert bool(((f[i] - c[i]).abs() <= 1e-14 * scale).all()), (tag, NAMES[i])
        assert any(bool(f[i][-1].abs().max() > 0) for i in range(4)), tag   # the last node written
        del c, f
        torch.cuda.synchronize()


def test_round_trip_field_size(gpu_ctx):
    """rnlm -> nlm -> rnlm is the identity, and the rebuilt nlm satisfies
    n_l^{-m} = (-1)^m conj(n_l^m), over 2^18 nodes at L = 12"""
    import torch
    from specfab_amd import batch as B
    L, P = 12, 1 << 18
    g = torch.Generator(device="cuda").manual_seed(3)
    r = torch.complex(torch.randn(P, O.rnlm_len(L), dtype=torch.float64, device="cuda", generator=g),
                      torch.randn(P, O.rnlm_len(L), dtype=torch.float64, device="cuda", generator=g))
    z = B.rnlm_to_nlm_batch(L, r)
    assert torch.equal(B.nlm_to_rnlm_batch(L, z), r)
    lm = O.lm_list(L)
    pos = {x: k for k, x in enumerate(lm)}
    neg = [k for k, (l, m) in enumerate(lm) if m < 0]
    mir = [pos[(l, -m)] for (l, m) in (lm[k] for k in neg)]
    sgn = torch.tensor([(-1.0) ** lm[k][1] for k in neg], dtype=torch.float64, device="cuda")
    assert torch.equal(z[:, neg], sgn * torch.conj(z[:, mir]))


def test_argument_errors(gpu_ctx):
    import torch
    from specfab_amd import _lib, batch as B
    with pytest.raises(Exception):
        B.nlm_to_rnlm_batch(7, torch.zeros((2, 36), dtype=torch.complex128, device="cuda"))
    assert _lib.lib().sfg_rnlm_len(32) < 0
    assert B.reduce_M_batch(4, torch.zeros((0, 15, 15), dtype=torch.complex128, device="cuda"))[0].shape == (0, 9, 9)


def test_M_REG_batch_vs_reference(gpu_ctx):
    """M_REG(D) of many nodes in one call (the per-node assembly of
    src/specfabpy/fenics/CPO.py:202) against the reference's M_REG of the
    ops fixtures (tests/golden/ops_L{4,8,12}.npz) at 1e-14, each node its own
    D, and reduced to the 2-D form like M_LROT and M_DDRX_src."""
    import torch
    from specfab_amd import batch as B
    for L in (4, 8, 12):
        g = load_golden("ops_L%d.npz" % L)
        D = torch.from_numpy(np.concatenate([g["eps"], 2.0 * g["eps"]])).cuda()
        M = B.M_REG_batch(L, D).cpu().numpy()
        K = len(g["eps"])
        for k in range(K):
            assert relerr(M[k], g["M_REG"][k]) < 1e-14
            assert relerr(M[K + k], 2.0 * g["M_REG"][k]) < 1e-14
        red = B.reduce_M_batch(L, torch.from_numpy(M.astype(np.complex128)).cuda())
        want = O.reduce_M(M.astype(np.complex128), L)
        for i, name in enumerate(NAMES):
            assert np.array_equal(red[i].cpu().numpy(), want[i]), name
