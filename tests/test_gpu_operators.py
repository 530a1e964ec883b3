"""GPU parity of the drop-in single-parcel API (specfab_amd.specfabpy, C-ABI
sfg_M_*, sfg_a2, sfg_ev_D2) and the batched moments against the reference's
golden operators (tests/golden/ops_L*.npz)."""
import numpy as np
import pytest

from conftest import load_golden, relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L", [4, 8, 12])
def test_operators_vs_reference(gpu_ctx, L):
    from specfab_amd import specfabpy as sf
    g = load_golden("ops_L%d.npz" % L)
    lm, n = sf.init(L)
    assert n == g["M_CDRX"].shape[0] and lm.shape == (2, n)
    for k in range(len(g["iota"])):
        nlm = g["nlm"][k]
        assert relerr(sf.M_LROT(nlm, g["eps"][k], g["omg"][k], g["iota"][k], g["zeta"][k]), g["M_LROT"][k]) < 1e-13
        assert relerr(sf.M_DDRX_src(nlm, g["tau"][k]), g["M_DDRX_src"][k]) < 1e-13
        assert relerr(sf.M_DDRX(nlm, g["tau"][k]), g["M_DDRX"][k]) < 1e-13
        assert relerr(sf.M_REG(nlm, g["eps"][k]), g["M_REG"][k]) < 1e-14
        assert relerr(sf.a2(nlm), g["a2"][k]) < 1e-14
        assert abs(sf.ev_D2(nlm, g["tau"][k]) - g["ev_D2"][k]) < 1e-13 * abs(g["ev_D2"][k])
    assert relerr(sf.M_CDRX(g["nlm"][0]).real, g["M_CDRX"]) == 0.0


@pytest.mark.parametrize("L", [4, 8, 12])
def test_batched_moments(gpu_ctx, L):
    from specfab_amd.batch import FabricBatch
    g = load_golden("ops_L%d.npz" % L)
    b = FabricBatch(L, nlm=g["nlm"], ugrad=np.zeros((3, 3, 3)))
    assert relerr(b.a2().cpu().numpy(), g["a2"]) < 1e-14
    assert relerr(b.a4_mandel().cpu().numpy(), g["a4v"]) < 1e-13
