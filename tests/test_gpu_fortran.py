"""Fortran host code -> ISO_C_BINDING module specfab_gpu -> libspecfab_gpu:
the one-parcel L=8 uniaxial-compression run (LROT + REG, Euler dt=1, 1000
steps) driven from Fortran (tests/fortran/c1_parcel.f90) with both the
single-parcel operators and the batched stepper, against the reference's own
trajectory (tests/golden/c1_traj.npz)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden, relerr

pytestmark = pytest.mark.gpu
EXE = os.path.join(ROOT, "specfab_amd", "fortran", "_build", "c1_parcel")
FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")


def _need(exe):
    """The Fortran test programs are built by __graft_entry__.build() with
    amdflang: without the compiler the face cannot be built (skip); with it,
    a missing program is a build failure."""
    if not os.path.exists(exe):
        if not os.path.exists(FC):
            pytest.skip("amdflang not present: the Fortran face cannot be built here")
        pytest.fail("%s missing although amdflang is present: run the build" % exe)


def test_fortran_c1_parcel():
    _need(EXE)
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = [l.split() for l in out.stdout.splitlines() if l.startswith("nlm ")]
    host = np.array([float(r[2]) + 1j * float(r[3]) for r in rows])
    batched = np.array([float(r[4]) + 1j * float(r[5]) for r in rows])
    ref = load_golden("c1_traj.npz")["traj"][-1]
    assert relerr(host, ref) < 1e-10
    assert relerr(batched, ref) < 1e-10


EXE2 = os.path.join(ROOT, "specfab_amd", "fortran", "_build", "homog_faces")


def test_fortran_homog_faces(tmp_path):
    """Fortran host code calling a4, Vi_elastic_tranisotropic (Hill),
    Vi_electromagnetic_tranisotropic, a2_to_nlm, a4_to_nlm, Gamma0 and
    apply_bounds of the specfab_gpu module, against the reference fixtures
    (wave_L8, init_L8); 1e-10 relative (1e-13 for a4)."""
    _need(EXE2)
    w = load_golden("wave_L8.npz")
    g = load_golden("init_L8.npz")
    h = load_golden("homog_L8.npz")
    o8 = load_golden("ops_L8.npz")
    ns, nd = 4, len(w["theta"])
    kb = next(k for k, x in enumerate(g["b_nlm"]) if not np.array_equal(x, g["b_out"][k]))
    f = tmp_path / "in.txt"
    fmt = lambda a: " ".join(repr(float(v)) for v in np.ravel(a))
    lines = ["%d %d" % (ns, nd)]
    for k in range(ns):
        lines += [fmt(w["nlm"][k].real), fmt(w["nlm"][k].imag)]
    lines += [fmt(w["theta"]), fmt(w["phi"]),
              fmt(np.concatenate([w["Lame"], [w["rho"], w["eps_m"], w["eps_t"], w["mu"]]]))]
    lines += [fmt(o8["tau"][0])]
    lines += [fmt(g["a2"][0]), fmt(np.ravel(g["a4"][0], order="F")), fmt(g["g_eps"][0]),
              fmt([g["g_T"][0], g["g_A"], g["g_Q"]]), fmt(g["b_nlm"][kb].real), fmt(g["b_nlm"][kb].imag)]
    f.write_text("\n".join(lines) + "\n")
    out = subprocess.run([EXE2, str(f)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = [l.split() for l in out.stdout.splitlines()]
    a4 = np.zeros((ns, 3, 3, 3, 3))
    vi = np.zeros((ns, nd, 3))
    ve = np.zeros((ns, nd, 2))
    a2n = np.zeros(6, complex)
    a4n = np.zeros(15, complex)
    bnd = np.zeros(45, complex)
    ebe = np.zeros((ns, 6))
    ebx = np.zeros((ns, 6))
    vb = np.zeros((ns, nd, 3))
    mds = np.zeros((45, 45), complex)
    G = None
    for r in rows:
        if r[0] == "ebe":
            ebe[int(r[1]) - 1] = [float(v) for v in r[2:8]]
        elif r[0] == "ebx":
            ebx[int(r[1]) - 1] = [float(v) for v in r[2:8]]
        elif r[0] == "vb":
            vb[int(r[1]) - 1, int(r[2]) - 1] = [float(v) for v in r[3:6]]
        elif r[0] == "mds":
            mds[int(r[1]) - 1, int(r[2]) - 1] = float(r[3]) + 1j * float(r[4])
        if r[0] == "a4":
            k, i, j, m, q = (int(v) - 1 for v in r[1:6])
            a4[k, i, j, m, q] = float(r[6])
        elif r[0] == "vi":
            vi[int(r[1]) - 1, int(r[2]) - 1] = [float(v) for v in r[3:6]]
        elif r[0] == "ve":
            ve[int(r[1]) - 1, int(r[2]) - 1] = [float(v) for v in r[3:5]]
        elif r[0] == "a2n":
            a2n[int(r[1]) - 1] = float(r[2]) + 1j * float(r[3])
        elif r[0] == "a4n":
            a4n[int(r[1]) - 1] = float(r[2]) + 1j * float(r[3])
        elif r[0] == "bnd":
            bnd[int(r[1]) - 1] = float(r[2]) + 1j * float(r[3])
        elif r[0] == "g0":
            G = float(r[1])
    assert relerr(a4, w["a4"][:ns]) < 1e-13
    assert relerr(vi, np.transpose(w["Vi_hill"][:ns], (0, 2, 1))) < 1e-10
    assert relerr(ve, np.transpose(w["Vi_em"][:ns], (0, 2, 1))) < 1e-10
    # the first wave_L8 states are the homog_L8 states
    assert np.array_equal(w["nlm"][:ns], h["nlm"][:ns])
    assert relerr(ebx, h["Eij_xyz_lin"][:ns]) < 1e-10
    assert relerr(ebe, h["Eij_eig_lin"][:ns]) < 1e-10
    assert relerr(vb, np.transpose(w["Vi_hill"][:ns], (0, 2, 1))) < 1e-10
    assert relerr(mds, o8["M_DDRX_src"][0]) < 1e-10
    assert relerr(a2n, g["a2_nlm"][0]) < 1e-14
    assert relerr(a4n, g["a4_nlm"][0]) < 1e-14
    assert abs(G - g["g_G"][0]) < 1e-14 * g["g_G"][0]
    assert relerr(bnd, g["b_out"][kb]) < 1e-14


EXE3 = os.path.join(ROOT, "specfab_amd", "fortran", "_build", "discrete_faces")
EXE4 = os.path.join(ROOT, "specfab_amd", "fortran", "_build", "ortho_faces")
EXE5 = os.path.join(ROOT, "specfab_amd", "fortran", "_build", "faces")


def test_fortran_discrete_faces(tmp_path):
    """Fortran host code calling ri_LROT and ri_to_nlm of the specfab_gpu
    module with the reference's argument shapes (ri0(N,3), D(Nt,3,3),
    ri(Nt,N,3)), against the reference fixture discrete.npz: 1e-13."""
    _need(EXE3)
    g = load_golden("discrete.npz")
    fmt = lambda a: " ".join(repr(float(v)) for v in np.ravel(a))
    N, Nt = g["lrot_ri0"].shape[0], g["lrot_ri"].shape[1]
    lines = ["%d %d %r %r" % (N, Nt, float(g["lrot_dt"]), float(g["lrot_iota"][0]))]
    lines += [fmt(r) for r in g["lrot_ri0"]]
    lines += [fmt(d) for d in g["lrot_D"]] + [fmt(w) for w in g["lrot_W"]]
    e, L = 1, 8
    lines += ["%d %d" % (len(g["r2n_wi"][e]), L)] + [fmt(r) for r in g["r2n_ri"][e]] + [fmt(g["r2n_wi"][e])]
    f = tmp_path / "in.txt"
    f.write_text("\n".join(lines) + "\n")
    out = subprocess.run([EXE3, str(f)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = [l.split() for l in out.stdout.splitlines()]
    ri = np.array([[float(x) for x in r[2:5]] for r in rows if r[0] == "ri"])
    nlm = np.array([float(r[2]) + 1j * float(r[3]) for r in rows if r[0] == "nlm"])
    assert np.abs(ri - g["lrot_ri"][0, -1]).max() < 1e-13
    assert relerr(nlm, g["r2n_nlm%d" % L][e]) < 1e-13


def test_fortran_ortho_faces(tmp_path):
    """Fortran host code calling Eij_orthotropic, Vi_elastic_orthotropic,
    Eij_orthotropic_discrete (bi(N,3)), Vi_elastic_orthotropic__discrete
    (mi(3,3,N)) and Cij_to_Lame_orthotropic of the specfab_gpu module, against
    the reference fixtures ortho_L4.npz (v given, Cartesian frame; 2e-6 /
    1e-7) and ortho_discrete.npz (ensemble 0; 1e-7 / 1e-11)."""
    _need(EXE4)
    g = load_golden("ortho_L4.npz")
    d = load_golden("ortho_discrete.npz")
    fmt = lambda a: " ".join(repr(float(v)) for v in np.ravel(a))
    Q = g["q"]
    lines = ["%d %d" % (len(Q), len(g["theta"])), fmt(g["Eg"][1]), fmt(g["Lame"]) + " " + repr(float(g["rho"])),
             fmt(g["theta"]), fmt(g["phi"])]
    for q in Q:
        for j in range(3):
            lines.append(fmt(np.stack([q[j].real, q[j].imag], 1)))
    T = d["grains_0"]
    lines.append("%d" % len(T))
    lines += [fmt(T[i]) for i in range(len(T))]
    lines.append(fmt(g["Cij"]))
    f = tmp_path / "in.txt"
    f.write_text("\n".join(lines) + "\n")
    out = subprocess.run([EXE4, str(f)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = [l.split() for l in out.stdout.splitlines()]
    E = np.array([[float(x) for x in r[2:8]] for r in rows if r[0] == "E"])
    V = np.array([[float(x) for x in r[3:6]] for r in rows if r[0] == "V"]).reshape(len(Q), -1, 3)
    ED = np.array([float(x) for x in next(r for r in rows if r[0] == "ED")[1:7]])
    VD = np.array([[float(x) for x in r[2:5]] for r in rows if r[0] == "VD"])
    lame = np.array([float(x) for x in next(r for r in rows if r[0] == "LAME")[1:10]])
    assert relerr(E, g["Eij_xyz_given_1"]) < 2e-6
    assert np.abs(V - np.transpose(g["Vi_given"], (0, 2, 1))).max() / V.max() < 1e-7
    assert relerr(ED, d["Eij_xyz_1"][0]) < 1e-7
    assert np.abs(VD - d["Vi"][0].T).max() / VD.max() < 1e-11
    assert relerr(lame, g["Lame"]) < 1e-15


def test_fortran_evw_eig3_qnorm_a6(tmp_path):
    """Fortran host code calling Evw_tranisotropic, Evw_orthotropic, eig3,
    Qnorm_tranisotropic and a6 of the specfab_gpu module (the reference
    module's argument orders and shapes) against the reference fixture
    faces_L8.npz (Evw 1e-10, orthotropic 2e-6, Qnorm 1e-10, eig3 1e-14) and
    the exact sixth moments of oracle/nonlinear.py (1e-14)."""
    _need(EXE5)
    from oracle import nonlinear as N
    g = load_golden("faces_L8.npz")
    fmt = lambda a: " ".join(repr(float(v)) for v in np.ravel(a))
    X, V, Wv, T = g["evw_nlm"][:3], g["evw_v"], g["evw_w"], g["evw_tau"]
    S = len(V)
    lines = ["%d %d" % (len(X), S), fmt(g["qn_Lame"]), fmt(g["evwo_Eg"])]
    state = lambda s_: fmt(np.concatenate([V[s_], Wv[s_], T[s_].ravel()]))
    for x in X:
        lines.append(fmt(np.stack([x.real, x.imag], 1)))
        lines += [state(s_) for s_ in range(S)]
    for j in range(3):
        q = g["evwo_q"][0][j]
        lines.append(fmt(np.stack([q.real, q.imag], 1)))
    lines += [state(s_) for s_ in range(S)]
    lines.append(fmt(g["eig_M"][0]))
    f = tmp_path / "in.txt"
    f.write_text("\n".join(lines) + "\n")
    out = subprocess.run([EXE5, str(f)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = [l.split() for l in out.stdout.splitlines()]
    Q = np.array([[float(v) for v in r[2:11]] for r in rows if r[0] == "Q"]).reshape(-1, 3, 3)
    A6 = np.array([[float(v) for v in r[2:6]] for r in rows if r[0] == "A6"])
    E = np.array([float(r[3]) for r in rows if r[0] == "EVW"]).reshape(len(X), S)
    EO = np.array([float(r[2]) for r in rows if r[0] == "EVWO"])
    eig = np.array([float(v) for v in next(r for r in rows if r[0] == "EIG")[1:13]])
    assert relerr(Q, g["qn_hill"][:3]) < 1e-10
    assert np.all(np.abs(E - g["evw_1_0.0125_1"][:3]) <= 1e-10 * np.maximum(1, np.abs(E)))
    assert np.all(np.abs(EO - g["evwo_given"][0]) <= 2e-6 * np.maximum(1, np.abs(EO)))
    for k, x in enumerate(X):
        m6 = N.moments(x, (6,))[6]
        want = [m6[0, 0, 0, 0, 0, 0], m6[0, 0, 1, 1, 2, 2], m6[0, 1, 2, 2, 2, 2], m6[2, 2, 2, 2, 2, 2]]
        assert np.abs(A6[k] - want).max() < 1e-14
    lam, e = eig[:3], eig[3:].reshape(3, 3)
    assert np.abs(lam - g["eig3_lami"][0]).max() < 1e-14
    for i in range(3):
        assert min(np.abs(e[i] - g["eig3_ei"][0][i]).max(), np.abs(e[i] + g["eig3_ei"][0][i]).max()) < 1e-12
