"""Pins for the initial-state, rate-factor, bounds and integrator fixtures
(tests/golden/init_L8.npz, written from the reference's own Fortran by
tests/golden/make_golden.py) and for the host-side kinematics of
specfab_amd.specfabpy:

  * a2_to_nlm / a4_to_nlm invert a2 / a4 (round trip through the reference's
    a2 and a4 to float32-literal accuracy) and are exact on the isotropic state;
  * Gamma0 equals A sqrt(tr(eps.eps)/2) exp(-Q/(R T)) with the gas constant
    read as a float32 literal (and measurably not with the FP64 value);
  * apply_bounds leaves states within the delta-function bound unchanged and
    scales the others onto it;
  * the integrator trajectories conserve n00 to float32 storage accuracy and
    the kinematics (pure / simple shear F, ugrad, strain -> time) match the
    reference's deformationmodes, which the fixture's F arrays hold.
With the reference library present (oracle/_ref), the fixture is also
regenerated and compared bit for bit."""
import numpy as np
import pytest

from conftest import load_golden


def _ref():
    from oracle import ref
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    r = ref.RefLib()
    r.init(8)
    return r


def _S(nlm, l):
    ls = np.concatenate([np.full(2 * k + 1, k) for k in range(0, 9, 2)])
    return np.sum(np.abs(nlm[ls == l]) ** 2) / (2 * l + 1)


def test_a2_a4_to_nlm_invert_moments():
    g = load_golden("init_L8.npz")
    r = _ref()
    for a2, z in zip(g["a2"], g["a2_nlm"]):
        x = np.zeros(45, dtype=np.complex128)
        x[:6] = z
        assert np.abs(r.a2(x) - a2).max() < 1e-6
    for a4, z in zip(g["a4"], g["a4_nlm"]):
        x = np.zeros(45, dtype=np.complex128)
        x[:15] = z
        assert np.abs(r.a4(x) - a4).max() < 1e-6
    iso = r.a2_to_nlm(np.eye(3) / 3)
    # isotropic a2: n00 = 1/sqrt(4 pi), the rest zero up to the float32 literals (~2e-8)
    assert abs(iso[0] - 1 / np.sqrt(4 * np.pi)) < 1e-15 and np.abs(iso[1:]).max() < 1e-7


def test_gamma0_float32_gas_constant():
    g = load_golden("init_L8.npz")
    R32, R64 = float(np.float32(8.31446261815324)), 8.31446261815324
    for e, T, G in zip(g["g_eps"], g["g_T"], g["g_G"]):
        I2 = np.sqrt(0.5 * np.trace(e @ e))
        assert abs(G - g["g_A"] * I2 * np.exp(-g["g_Q"] / (R32 * T))) < 1e-14 * G
        assert abs(G - g["g_A"] * I2 * np.exp(-g["g_Q"] / (R64 * T))) > 1e-9 * G


def test_apply_bounds_fixture():
    g = load_golden("init_L8.npz")
    for x, y in zip(g["b_nlm"], g["b_out"]):
        S0 = x[0].real ** 2
        for l, sl in ((2, slice(1, 6)), (4, slice(6, 15))):
            rel = _S(x, l) / S0
            if rel > 1:
                assert abs(_S(y, l) / S0 - 1) < 1e-13
                assert np.abs(y[sl] * np.sqrt(rel) - x[sl]).max() < 1e-15
            else:
                assert np.array_equal(y[sl], x[sl])
        assert np.array_equal(y[0], x[0]) and np.array_equal(y[15:], x[15:])
    assert any(_S(x, 2) / x[0].real ** 2 > 1 for x in g["b_nlm"])


def test_trajectories_and_kinematics():
    from specfab_amd import specfabpy as sf
    g = load_golden("init_L8.npz")
    for k in ("traj_ps", "traj_ss", "traj_ps_custom"):
        t = g[k]
        assert t.dtype == np.complex64 and t.shape == (101, 45)
        assert np.abs(t[:, 0] - t[0, 0]).max() < 1e-4 * abs(t[0, 0])
        assert np.abs(t[-1, 1:]).max() > 1e-3          # the fabric evolved
    dt_ps = float(g["ps_dt"])
    assert abs(sf.pureshear_strainii_to_t(-0.7, 1.0) / 100 - dt_ps) < 1e-16
    assert abs(sf.simpleshear_gamma_to_t(np.deg2rad(70), 1.0) / 100 - float(g["ss_dt"])) < 1e-16
    for k in range(101):
        assert np.abs(sf.pureshear_F(2, 0.0, 1.0, dt_ps * k) - g["ps_F"][k]).max() < 1e-15
        assert np.abs(sf.simpleshear_F(1, 1.0, float(g["ss_dt"]) * k) - g["ss_F"][k]).max() < 1e-14
    r = _ref()
    for ax in (0, 1, 2):
        assert np.abs(sf.pureshear_ugrad(ax, 0.3, 2.0) - r.pureshear_ugrad(ax, 0.3, 2.0)).max() == 0
        assert np.abs(sf.pureshear_F(ax, 0.3, 2.0, 0.7) - r.pureshear_F(ax, 0.3, 2.0, 0.7)).max() < 1e-15
        assert np.abs(sf.simpleshear_ugrad(ax, 2.0) - r.simpleshear_ugrad(ax, 2.0)).max() == 0
        assert np.abs(sf.simpleshear_F(ax, 2.0, 0.7) - r.simpleshear_F(ax, 2.0, 0.7)).max() < 1e-15


def test_fixture_regenerates():
    r = _ref()
    import importlib.util
    import os
    spec = importlib.util.spec_from_file_location(
        "make_golden", os.path.join(os.path.dirname(__file__), "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    rec = mg.init_fixture(r)
    g = load_golden("init_L8.npz")
    for k, v in rec.items():
        assert np.array_equal(np.asarray(v), g[k]), k
