"""Pins for the discrete-grain oracle (oracle/discrete.py), the generated a6
map (specfab_amd/tables/make_a6_map.py) and the reference fixture
tests/golden/discrete.npz (written from the reference's own Fortran by
tests/golden/make_golden.py):

  * ri_LROT: the numpy restatement equals the reference fixture; a pure spin
    (iota = 0) about z turns every axis by exactly atan(dt w) per Euler step
    and tilts it off z by exactly sqrt(1 + (dt w)^2) in tan(theta) (the
    renormalised step of I + dt W); uniaxial compression along z with
    iota = 1 (c-axes) turns axes towards z and with iota = -1 away from it;
    the trajectories stay unit vectors;
  * ri_to_nlm: the restatement's a6 map (a least-squares fit of scipy's
    spherical harmonics) equals the reference; the sympy-derived map in the
    generated header equals both; a single grain's expansion is conj(Y_lm) of
    its axis (FP64 rows exactly, REAL(4)-root rows to float32 accuracy); the
    12 vertices of an icosahedron (a spherical 5-design) have zero l = 2, 4
    coefficients and non-zero l = 6 ones; the l <= 2 / l <= 4 rows of the a6
    path agree with the reference's a2 / a4 paths (which differ from it only
    by their own REAL(4) literals) for unit axes and unit total weight.
With the reference library present (oracle/_ref), the fixture is also
regenerated and compared bit for bit."""
import os
import re

import numpy as np
import pytest
from scipy.special import sph_harm_y

from conftest import ROOT, load_golden
from oracle import discrete as O


def test_ri_LROT_restatement_matches_reference():
    g = load_golden("discrete.npz")
    for j, io in enumerate(g["lrot_iota"]):
        ri = O.ri_LROT(g["lrot_ri0"], float(g["lrot_dt"]), g["lrot_ri"].shape[1], g["lrot_D"], g["lrot_W"], io)
        assert np.abs(ri - g["lrot_ri"][j]).max() < 1e-13
    assert np.abs(np.linalg.norm(g["lrot_ri"], axis=-1) - 1).max() < 1e-14


def test_ri_LROT_spin_closed_form():
    rng = np.random.default_rng(3)
    N, Nt, dt, w = 16, 30, 0.07, 1.3
    phi0 = rng.uniform(0, 2 * np.pi, N)
    th = rng.uniform(0.2, 2.9, N)
    r0 = np.stack([np.sin(th) * np.cos(phi0), np.sin(th) * np.sin(phi0), np.cos(th)], axis=1)
    Wm = np.zeros((Nt, 3, 3))
    Wm[:, 0, 1], Wm[:, 1, 0] = -w, w            # dr/dt = W r: counter-clockwise about z
    ri = O.ri_LROT(r0, dt, Nt, np.zeros((Nt, 3, 3)), Wm, 0.0)
    k = np.arange(Nt)[:, None]
    phi = np.arctan2(ri[..., 1], ri[..., 0])
    expect = phi0[None, :] + k * np.arctan(dt * w)
    assert np.abs(np.angle(np.exp(1j * (phi - expect)))).max() < 1e-12
    # (I + dt W) stretches the xy part by sqrt(1 + (dt w)^2) and keeps z:
    # tan(theta) grows by that factor per step
    tan_th = np.hypot(ri[..., 0], ri[..., 1]) / ri[..., 2]
    expect_tan = np.tan(th)[None, :] * (1 + (dt * w) ** 2) ** (k / 2)
    assert np.abs(tan_th / expect_tan - 1).max() < 1e-12


def test_ri_LROT_compression_direction():
    rng = np.random.default_rng(4)
    r0 = rng.normal(size=(200, 3))
    r0 /= np.linalg.norm(r0, axis=1, keepdims=True)
    Nt = 200
    D = np.tile(np.diag([0.5, 0.5, -1.0]), (Nt, 1, 1))
    Wm = np.zeros((Nt, 3, 3))
    c = O.ri_LROT(r0, 0.05, Nt, D, Wm, 1.0)
    a = O.ri_LROT(r0, 0.05, Nt, D, Wm, -1.0)
    assert np.mean(np.abs(c[-1, :, 2])) > 0.99            # c-axes gather at the compression axis
    assert np.mean(np.abs(a[-1, :, 2])) < 0.05            # iota = -1 sends them to the plane


def test_ri_to_nlm_restatement_matches_reference():
    g = load_golden("discrete.npz")
    for L in (6, 8, 10):
        for e in range(len(g["r2n_ri"])):
            z = O.ri_to_nlm(g["r2n_ri"][e], g["r2n_wi"][e], L)
            ref = g["r2n_nlm%d" % L][e]
            assert np.abs(z - ref).max() / np.abs(ref).max() < 1e-14
            assert np.all(ref[28:] == 0)
    for r, ref in zip(g["r2n_one_ri"], g["r2n_one_nlm"]):
        assert np.abs(O.ri_to_nlm(r[None, :], np.ones(1), 6) - ref).max() < 1e-14


def test_single_grain_is_conj_Ylm():
    g = load_golden("discrete.npz")
    for r, ref in zip(g["r2n_one_ri"], g["r2n_one_nlm"]):
        th, ph = np.arccos(r[2]), np.arctan2(r[1], r[0])
        Y = np.array([np.conj(sph_harm_y(l, m, th, ph)) for l, m in O.lm6()])
        half = np.array([(l, abs(m)) in O.HALF_ROOT for l, m in O.lm6()])
        assert np.abs(ref[~half] - Y[~half]).max() < 1e-13
        assert np.abs(ref[half] - Y[half]).max() < 1e-7
        assert np.abs(ref[half] - Y[half]).max() > 1e-12      # the REAL(4) roots are visible


def test_icosahedron_is_isotropic_to_l4():
    p = (1 + np.sqrt(5)) / 2
    v = np.array([[0, s1, s2 * p] for s1 in (1, -1) for s2 in (1, -1)], dtype=float)
    V = np.concatenate([v, v[:, [1, 2, 0]], v[:, [2, 0, 1]]])
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    z = O.ri_to_nlm(V, np.full(12, 1 / 12), 6)
    assert abs(z[0] - 1 / (2 * np.sqrt(np.pi))) < 1e-15
    assert np.abs(z[1:15]).max() < 1e-8            # REAL(4) roots leave 1e-8-level residues
    assert np.abs(z[15:28]).max() > 0.05


def test_a6_path_agrees_with_a2_a4_paths():
    g = load_golden("discrete.npz")
    for e in range(len(g["r2n_ri"])):
        z6 = g["r2n_nlm6"][e]
        assert np.abs(g["r2n_nlm2"][e] - z6[:6]).max() < 1e-6
        assert np.abs(g["r2n_nlm4"][e] - z6[:15]).max() < 1e-6


def _generated_map():
    path = os.path.join(ROOT, "specfab_amd", "csrc", "generated", "sfg_a6_map.h")
    if not os.path.exists(path):
        pytest.skip("generated a6 map missing: run __graft_entry__.build()")
    txt = open(path).read()
    kind = np.array([int(x) for x in re.search(r"A6_KIND\[28\] = \{([^}]*)\}", txt).group(1).split(",")])
    q = np.array([float(x) for x in re.search(r"A6_Q\[28\] = \{([^}]*)\}", txt).group(1).split(",")])
    rows = re.findall(r"^\s*\{([^}]*)\},\s*// (re|im) n_(\d)(\d)", txt, flags=re.M)
    W = np.array([[float.fromhex(x.strip()) for x in r[0].split(",")] for r in rows])
    tags = [(r[1], int(r[2]), int(r[3])) for r in rows]
    mons = [tuple(int(c) for c in m.split(",")) for m in re.findall(r"\{(\d, \d, \d)\}", txt)[:28]]
    return kind, q, W, tags, mons


def _eval_generated(ri, wi, L):
    """The generated header's map evaluated as the kernel does, in numpy."""
    kind, q, W, tags, mons = _generated_map()
    r = np.asarray(ri)
    mom = np.array([np.sum(wi * r[:, 0] ** a * r[:, 1] ** b * r[:, 2] ** c) for a, b, c in mons])
    K = np.where(kind == 0, 1 / np.sqrt(np.pi),
                 np.where(kind == 1, np.sqrt(q / np.pi),
                          np.where(kind == 2, np.sqrt(q.astype(np.float32)).astype(np.float64) / np.sqrt(np.pi),
                                   (np.float32(3) * np.sqrt(q.astype(np.float32))).astype(np.float64) / 3
                                   / np.sqrt(np.pi))))
    real = K * (W @ mom)
    out = np.zeros((L + 1) * (L + 2) // 2, dtype=np.complex128)
    pos = {lm: i for i, lm in enumerate(O.lm6())}
    for (tag, l, m), v in zip(tags, real):
        if tag == "re":
            out[pos[(l, m)]] += v
        else:
            out[pos[(l, m)]] += 1j * v
    for l in (2, 4, 6):                       # n_l^-m = (-1)^m conj(n_l^m)
        for m in range(1, l + 1):
            out[pos[(l, -m)]] = (-1) ** m * np.conj(out[pos[(l, m)]])
    out[0] = 1 / (2 * np.sqrt(np.pi))
    return out


def test_generated_map_matches_reference():
    g = load_golden("discrete.npz")
    for e in range(len(g["r2n_ri"])):
        z = _eval_generated(g["r2n_ri"][e], g["r2n_wi"][e], 8)
        ref = g["r2n_nlm8"][e]
        assert np.abs(z - ref).max() / np.abs(ref).max() < 1e-14


def test_fixture_regenerates_bit_for_bit():
    from oracle import ref
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(ROOT, "tests", "golden",
                                                                              "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    new = mg.discrete_fixture(ref.RefLib())
    old = load_golden("discrete.npz")
    assert set(new) == set(old)
    for k in new:
        assert np.array_equal(np.asarray(new[k]), old[k]), k
