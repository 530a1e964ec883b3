"""CPU checks of the HIP stepping kernels: the parameter kernel, the generic
stepper and the generated register-resident steppers are compiled as plain
C++ against a host shim (specfab_amd/csrc/sfg_emul.h, -DSFG_HOST_EMUL) and run
lane by lane on the CPU, then compared with the C oracle.  This validates the
generated code's indexing and arithmetic before it runs on a GPU."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import ROOT, relerr

CSRC = os.path.join(ROOT, "specfab_amd", "csrc")
GEN = os.path.join(CSRC, "generated")
_cache = {}


def _block(src):
    """(threads per workgroup, parcels per workgroup) of a generated kernel."""
    import re
    text = open(src).read()
    bs = int(re.search(r"__launch_bounds__\((\d+),", text).group(1))
    m = re.search(r"hipLaunchKernelGGL\(\w+, dim3\(\(unsigned\)\(\(a\.P \+ \d+\) / (\d+)\)\)", text)
    return bs, int(m.group(1))


def emul_lib(L, prec, v=0):
    key = (L, prec, v)
    if key in _cache:
        return _cache[key]
    src = os.path.join(GEN, "sfg_fast_L%d_%s_v%d.hip" % (L, prec, v))
    if not os.path.exists(src):
        pytest.skip("generated kernels not present (run the build)")
    out = os.path.join(tempfile.gettempdir(), "sfg_emul_%d_%s_v%d_%d.so" % (L, prec, v, os.getpid()))
    cmd = ["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-DSFG_HOST_EMUL",
           "-DKERNEL_NAME=evolve_fast_L%d_%s_v%d" % (L, prec, v), "-DKERNEL_BLOCK=%d" % _block(src)[0],
           "-DKERNEL_PARCELS=%d" % _block(src)[1], "-pthread", "-I", CSRC, "-x", "c++",
           os.path.join(ROOT, "tests", "emul_driver.cpp"), os.path.join(CSRC, "sfg_kernels.hip"),
           os.path.join(CSRC, "sfg_basis.cpp"), src, "-o", out]
    subprocess.check_call(cmd)
    lib = ctypes.CDLL(out)
    d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    i = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    c = ctypes
    lib.emul_evolve.argtypes = [c.c_int, c.c_int64, c.c_int, c.c_int, c.c_int, c.c_double, c.c_double, c.c_double,
                                d, d, d, d, c.c_int, i]
    _cache[key] = lib
    return lib


def real_index(L):
    """(complex index, part) of each real-basis coefficient (include/specfab_gpu.h)."""
    out = []
    base = 0
    for l in range(0, L + 1, 2):
        for m in range(0, l + 1):
            out.append((base + l + m, 0))
            if m > 0:
                out.append((base + l + m, 1))
        base += 2 * l + 1
    return out


def to_state(nlm, L):
    ri = real_index(L)
    return np.ascontiguousarray(np.array([nlm[:, j].real if c == 0 else nlm[:, j].imag for (j, c) in ri]))


def to_nlm(x, L):
    lm = [(l, m) for l in range(0, L + 1, 2) for m in range(-l, l + 1)]
    ri = {v: a for a, v in enumerate(real_index(L))}
    P = x.shape[1]
    out = np.zeros((P, len(lm)), complex)
    for j, (l, m) in enumerate(lm):
        jp = lm.index((l, abs(m)))
        v = x[ri[(jp, 0)]] + (1j * x[ri[(jp, 1)]] if m != 0 else 0)
        out[:, j] = v if m >= 0 else (-1) ** m * np.conj(v)
    return out


PROCS = {"LROT": 1, "DDRX": 2, "CDRX": 4, "REG": 8}


def _variants(L, prec):
    return sorted(int(f.split("_v")[1].split(".")[0]) for f in os.listdir(GEN)
                  if f.startswith("sfg_fast_L%d_%s_v" % (L, prec))) if os.path.isdir(GEN) else [0]


# every variant of the tuned L = 8 kernels and of the smallest window-stepper L
# per precision (14 FP64, 18 FP32); the first variant elsewhere
CASES = [(L, prec, v) for (L, prec) in [(4, "f64"), (8, "f64"), (12, "f64"), (14, "f64"), (8, "f32"), (12, "f32"),
                                        (18, "f32")]
         for v in (_variants(L, prec) if (L, prec) in ((8, "f64"), (8, "f32"), (14, "f64"), (18, "f32")) else [0])]


@pytest.mark.parametrize("L,prec,v", CASES)
@pytest.mark.parametrize("scheme", ["rk4", "euler"])
def test_generated_kernel_vs_oracle(oracle, L, prec, v, scheme):
    from specfab_amd import workloads as W
    rng = np.random.default_rng(L * 10 + (scheme == "rk4"))
    P = 70   # one full and one partial wavefront
    nlm0 = W.perturbed_isotropic(rng, P, L, sigma=1e-2)
    ug = np.ascontiguousarray(W.traceless_gradients(rng, P))
    g0 = 10.0 ** rng.uniform(-4, -2, P)
    lam = 10.0 ** rng.uniform(-4, -3, P)
    procs = ("LROT", "REG", "DDRX", "CDRX")
    mask = sum(PROCS[p] for p in procs)
    steps, dt = 6, 0.5
    ref = oracle.evolve(L, nlm0, ug, g0, lam, steps, dt, scheme, procs)
    lib = emul_lib(L, prec, v)
    tol = 1e-12 if prec == "f64" else 2e-6
    for which in (0, 1) if prec == "f64" else (0,):
        x = to_state(nlm0, L)
        st = np.zeros(P, np.int32)
        lib.emul_evolve(L, P, mask, 1 if scheme == "rk4" else 0, steps, dt, 1.0, 0.0, ug.reshape(-1).copy(), g0,
                        lam, x, which, st)
        out = to_nlm(x, L)
        assert relerr(out, ref) < tol, (which, relerr(out, ref))
        assert not st.any()


# the stepper families: literal-constant stage buffers (v0), register state
# with the rolled next-stage input (v5), split resident steppers with two and
# four waves per parcel group and workgroup barriers (v6, v7)
@pytest.mark.parametrize("L,v", [(8, 0), (8, 5), (8, 6), (8, 7)])
def test_emulation_under_sanitizers(L, v):
    """The parameter kernel, the generic stepper and a generated stepper,
    compiled for the host with AddressSanitizer and UndefinedBehaviorSanitizer
    (any report is fatal), step one full and one partial wavefront with RK4
    and Euler; the generated stepper agrees with the generic one."""
    src = os.path.join(GEN, "sfg_fast_L%d_f64_v%d.hip" % (L, v))
    if not os.path.exists(src):
        pytest.skip("generated kernels not present (run the build)")
    exe = os.path.join(tempfile.gettempdir(), "sfg_emul_san_%d_v%d_%d" % (L, v, os.getpid()))
    bs, parcels = _block(src)
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-DSFG_HOST_EMUL", "-DKERNEL_NAME=evolve_fast_L%d_f64_v%d" % (L, v),
           "-DKERNEL_BLOCK=%d" % bs, "-DKERNEL_PARCELS=%d" % parcels, "-pthread", "-I", CSRC, "-x", "c++",
           os.path.join(ROOT, "tests", "emul_driver.cpp"), os.path.join(ROOT, "tests", "emul_sanitize_main.cpp"),
           os.path.join(CSRC, "sfg_kernels.hip"), os.path.join(CSRC, "sfg_basis.cpp"), src, "-o", exe]
    subprocess.check_call(cmd)
    try:
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
        for scheme in (1, 0):
            r = subprocess.run([exe, str(L), str(scheme)], capture_output=True, text=True, timeout=600, env=env)
            print(r.stdout.strip())
            assert r.returncode == 0, r.stdout + r.stderr[-4000:]
            assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-4000:]
    finally:
        os.remove(exe)


_wave = {}


def emul_wave_lib():
    if "lib" in _wave:
        return _wave["lib"]
    out = os.path.join(tempfile.gettempdir(), "sfg_emul_wave_%d.so" % os.getpid())
    cmd = ["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-DSFG_HOST_EMUL", "-I", CSRC, "-x", "c++",
           os.path.join(ROOT, "tests", "emul_wave.cpp"), os.path.join(CSRC, "sfg_wave.hip"),
           os.path.join(CSRC, "sfg_wave_host.cpp"), os.path.join(CSRC, "sfg_basis.cpp"), "-o", out]
    if not os.path.exists(os.path.join(GEN, "..", "generated", "gaunt_L20.inc")):
        pytest.skip("generated tables not present (run the build)")
    subprocess.check_call(cmd)
    lib = ctypes.CDLL(out)
    d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    i = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    c = ctypes
    lib.emul_vi_elastic.argtypes = [c.c_int, c.c_int64, d, c.c_double, d, c.c_double, c.c_int, d, d, d, i]
    lib.emul_vi_em.argtypes = [c.c_int, c.c_int64, d, c.c_double, c.c_double, c.c_double, c.c_int, d, d, d]
    lib.emul_vi_elastic_k.argtypes = [c.c_int, c.c_int64, d, c.c_double, d, d, c.c_double, c.c_int, d, d, d, i]
    lib.emul_a4.argtypes = [c.c_int, c.c_int64, d, d]
    dn = ctypes.c_void_p
    lib.emul_eij_ortho.argtypes = [c.c_int, c.c_int64, d, d, dn, dn, d, d]
    lib.emul_vi_ortho.argtypes = [c.c_int, c.c_int64, d, d, dn, d, c.c_double, c.c_int, d, d, d]
    _wave["lib"] = lib
    return lib


def test_wave_kernels_vs_reference():
    """Elastic/EM velocity and full-a4 kernels (the 6x6 Mandel Reuss solve in
    place of the reference's 9x9 inverse) against the reference fixture."""
    from conftest import load_golden
    g = load_golden("wave_L8.npz")
    lib = emul_wave_lib()
    X = to_state(g["nlm"], 8)
    P = X.shape[1]
    th, ph = g["theta"], g["phi"]
    nd = len(th)
    for tag, alpha in (("reuss", 0.0), ("hill", 0.5), ("voigt", 1.0)):
        V = np.zeros((P, nd, 3))
        st = np.zeros(P, np.int32)
        lib.emul_vi_elastic(8, P, X, alpha, g["Lame"], float(g["rho"]), nd, th, ph, V, st)
        ref = np.transpose(g["Vi_" + tag], (0, 2, 1))
        assert relerr(V[:24], ref[:24]) < 1e-12, tag
        assert relerr(V[24:], ref[24:]) < 1e-7, tag
        assert not st.any()
    V = np.zeros((P, nd, 2))
    lib.emul_vi_em(8, P, X, float(g["eps_m"]), float(g["eps_t"]), float(g["mu"]), nd, th, ph, V)
    assert relerr(V, np.transpose(g["Vi_em"], (0, 2, 1))) < 1e-13
    A = np.zeros((P, 81))
    lib.emul_a4(8, P, X, A)
    assert relerr(A.reshape(P, 3, 3, 3, 3), g["a4"]) < 1e-14


def _ortho_ci(Eg):
    B = np.asarray(Eg, float) ** (2.0 / 2)
    lam = [-B[0] + B[1] + B[2], B[0] - B[1] + B[2], B[0] + B[1] - B[2], B[3], B[4], B[5]]
    return np.array([4.0 / 3 * v for v in lam[:3]] + [2.0 * v for v in lam[3:]])


def test_ortho_kernels_vs_reference():
    """Orthotropic Eij (pair-form Sachs tensor, Cartesian and a random frame,
    v given, derived, and mixed per parcel) and Vi (Voigt, six directions)
    kernels against the reference fixture (2e-6 / 1e-7: its REAL(4) joint
    bodies) and the numpy restatement (1e-11)."""
    from conftest import load_golden
    from oracle import orthotropic as O
    from oracle.oracle import Oracle
    orc = Oracle()
    g = load_golden("ortho_L4.npz")
    lib = emul_wave_lib()
    Q = g["q"]
    K = len(Q)
    X = [to_state(Q[:, i], 4) for i in range(3)]
    R = g["frame"]
    fr = np.ascontiguousarray(np.tile(R[None], (K, 1, 1)))
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None
    for gi, Eg in enumerate(g["Eg"]):
        ci = _ortho_ci(Eg)
        for tag, x3 in (("given", X[2]), ("derived", None)):
            for ftag, frames in (("xyz", None), ("rot", fr)):
                E = np.zeros((K, 6))
                lib.emul_eij_ortho(4, K, X[0], X[1], p(x3), p(frames), ci, E)
                assert relerr(E, g["Eij_%s_%s_%d" % (ftag, tag, gi)]) < 2e-6, (gi, tag, ftag)
                e = np.eye(3) if frames is None else R
                for k in range(K):
                    ora = O.Eij(Q[k, 0], Q[k, 1], Q[k, 2] if x3 is not None else 0 * Q[k, 2], e[0], e[1], e[2], Eg,
                                orc)
                    assert relerr(E[k], ora) < 1e-11, (gi, tag, ftag, k)
    # v given for some parcels and zero (derived) for others in one launch
    x3m = X[2].copy()
    x3m[:, 1] = 0.0
    E = np.zeros((K, 6))
    lib.emul_eij_ortho(4, K, X[0], X[1], p(x3m), None, _ortho_ci(g["Eg"][1]), E)
    assert relerr(E[[0, 2]], g["Eij_xyz_given_1"][[0, 2]]) < 2e-6
    assert relerr(E[1], g["Eij_xyz_derived_1"][1]) < 2e-6
    th, ph = g["theta"], g["phi"]
    nd = len(th)
    for tag, x3 in (("given", X[2]), ("derived", None)):
        V = np.zeros((K, nd, 3))
        lib.emul_vi_ortho(4, K, X[0], X[1], p(x3), g["Lame"], float(g["rho"]), nd, th, ph, V)
        ref = np.transpose(g["Vi_" + tag], (0, 2, 1))
        assert np.abs(V - ref).max() / ref.max() < 1e-7, tag
        for k in range(K):
            ora = O.Vi(Q[k, 0], Q[k, 1], Q[k, 2] if x3 is not None else None, g["Lame"], float(g["rho"]), th, ph,
                       orc)
            assert np.abs(V[k] - ora).max() / ora.max() < 1e-11, (tag, k)


def test_reuss_singular_branch():
    """The Reuss solve flags an exactly singular compliance (a zero LU pivot,
    the condition under which the reference's dgetrf returns info > 0 and it
    stops, src/wavepropagation.f90:347-352) with SFG_ST_REUSS_SING and NaN
    velocities.  Grain parameters from Cij_to_Lame never give a zero pivot
    (the xy-shear diagonal is at least 1 / (2 mu)), so the coefficients are
    given directly: with kf = (1, 0, 1/2, 0, 0) the xy-shear row of the
    Reuss matrix vanishes wherever a2_xy = 0 (no a4 or anticommutator term
    is left): for a state with only m = 0 coefficients also after its
    rotation to a direction (theta, phi), which leaves it mirror-symmetric
    about a plane through z.  The same states with the ice coefficients are regular, and
    the Voigt average (alpha = 1) does not use the Reuss solve."""
    lib = emul_wave_lib()
    L = 8
    n = 45
    X = np.zeros((n, 3))
    X[0] = 1 / np.sqrt(4 * np.pi)
    X[1, 1], X[6, 1] = 0.3, 0.2     # n_2^0, n_4^0: real basis rows of m = 0
    X[1, 2], X[6, 2] = -0.4, 0.1
    th, ph = np.array([0.0, 0.8]), np.array([0.0, 1.1])
    kf = np.array([1.0, 0.0, 0.5, 0.0, 0.0])
    kr = np.array([1.0, 2.0, 0.1, 0.2, 0.3])
    for alpha, sing in ((0.0, True), (0.5, True), (1.0, False)):
        V = np.zeros((3, 2, 3))
        st = np.zeros(3, np.int32)
        lib.emul_vi_elastic_k(L, 3, X, alpha, kf, kr, 917.0, 2, th, ph, V, st)
        if sing:
            assert (st == 8).all() and np.isnan(V).all(), (alpha, st)
        else:
            assert not st.any() and np.isfinite(V).all()
    from conftest import load_golden
    g = load_golden("wave_L8.npz")
    V = np.zeros((3, 2, 3))
    st = np.zeros(3, np.int32)
    lib.emul_vi_elastic(L, 3, X, 0.0, g["Lame"], 917.0, 2, th, ph, V, st)
    assert not st.any() and np.isfinite(V).all()


_rot = {}


def emul_rotate_lib():
    if "lib" in _rot:
        return _rot["lib"]
    out = os.path.join(tempfile.gettempdir(), "sfg_emul_rotate_%d.so" % os.getpid())
    cmd = ["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-DSFG_HOST_EMUL", "-I", CSRC, "-x", "c++",
           os.path.join(ROOT, "tests", "emul_rotate.cpp"), os.path.join(CSRC, "sfg_rotate.hip"),
           os.path.join(CSRC, "sfg_wave_host.cpp"), os.path.join(CSRC, "sfg_basis.cpp"), "-o", out]
    if not os.path.exists(os.path.join(GEN, "..", "generated", "gaunt_L20.inc")):
        pytest.skip("generated tables not present (run the build)")
    subprocess.check_call(cmd)
    lib = ctypes.CDLL(out)
    d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    i = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    c = ctypes
    lib.emul_rotate.argtypes = [c.c_int, c.c_int64, d, d, d, d]
    lib.emul_a4eig.argtypes = [c.c_int, c.c_int64, d, d, d, i]
    _rot["lib"] = lib
    return lib


def _rot_blocks(a, ref, L):
    """largest relative error per degree block of rotated states (P, n)"""
    out = {}
    for l in range(2, L + 1, 2):
        s = slice(l * (l - 1) // 2, (l + 1) * (l + 2) // 2)
        out[l] = float(np.abs(a[:, s] - ref[:, s]).max() / np.abs(ref[:, s]).max())
    return out


def test_rotate_kernel_vs_reference():
    """rotate_nlm (src/rotation.f90:29-55) on the CPU emulation of the
    rotation kernel, per-parcel angles, against the reference's outputs
    (tests/golden/rotate_L12.npz) for L = 4, 8, 12.  The l = 2 and l = 4
    blocks carry the reference's REAL(4) square roots of
    src/include/Dlmn.f90 and agree to 1e-12.  From l = 6 on, 1473 of 1630
    Wigner-D entries have the square-free part of their normalisation as a
    single REAL(4) factor (reproduced); the rest are sums of terms with
    different REAL(4) literals, so these blocks agree to 1e-7 (the float32
    rounding of the literals, measured <= 6.4e-8).  D(m, n) = exp(-i n phi)
    d_mn(theta): rotating by (0, phi) then (theta, 0) equals rotating by
    (theta, phi); a rotation preserves each block's power S(l) to 1e-7 (the
    REAL(4) normalisations make the reference's D^l unitary only to float32
    precision, and this reproduces them)."""
    from conftest import load_golden
    g = load_golden("rotate_L12.npz")
    lib = emul_rotate_lib()
    ang = g["angles"]
    for L in (4, 8, 12):
        X = g["nlm%d" % L]
        R = g["rot%d" % L]
        K, A = X.shape[0], len(ang)
        x = to_state(np.repeat(X, A, axis=0), L)
        th = np.tile(ang[:, 0], K).copy()
        ph = np.tile(ang[:, 1], K).copy()
        out = np.zeros_like(x)
        assert lib.emul_rotate(L, K * A, x, th, ph, out) == 0
        got = to_nlm(out, L)
        err = _rot_blocks(got, R.reshape(K * A, -1), L)
        print("L=%d rotate_nlm vs reference per degree:" % L, {l: "%.1e" % e for l, e in err.items()})
        for l, e in err.items():
            assert e < (1e-12 if l <= 4 else 1e-7), (L, l, e)
        # composition and power
        x2 = np.zeros_like(x)
        lib.emul_rotate(L, K * A, x, np.zeros_like(th), ph, x2)
        x3 = np.zeros_like(x)
        lib.emul_rotate(L, K * A, x2, th, np.zeros_like(ph), x3)
        assert np.abs(x3 - out).max() < 1e-13 * np.abs(out).max()
        Xr = np.repeat(X, A, axis=0)
        for l in range(2, L + 1, 2):
            s = slice(l * (l - 1) // 2, (l + 1) * (l + 2) // 2)
            p0 = (np.abs(Xr[:, s]) ** 2).sum(1)
            p1 = (np.abs(got[:, s]) ** 2).sum(1)
            assert np.abs(p1 - p0).max() < 1e-7 * p0.max()


def test_a4_eigentensors_and_isvalid_vs_reference():
    """a4_eigentensors (src/frames.f90:131-152) and nlm_isvalid
    (src/idealstate.f90:50-78) on the CPU emulation of their kernel against
    the reference (tests/golden/rotate_L12.npz): eigenvalues (ascending) to
    1e-12, eigentensors up to sign where the eigen-gap exceeds 1e-8, the
    eigentensors orthonormal in the Mandel inner product and
    sum_i ev_i Q_i (x) Q_i = a4; validity of the (nhat20, nhat40) grid equal to
    the reference's."""
    from conftest import load_golden
    g = load_golden("rotate_L12.npz")
    lib = emul_rotate_lib()
    X = g["a4eig_nlm"]
    P = X.shape[0]
    x = to_state(X, 12)
    Q = np.zeros((P, 6, 3, 3))
    ev = np.zeros((P, 6))
    ok = np.zeros(P, np.int32)
    lib.emul_a4eig(12, P, x, Q, ev, ok)
    assert np.abs(ev - g["a4eig_ev"]).max() < 1e-12
    gap = np.diff(g["a4eig_ev"], axis=1)
    for p in range(P):
        for i in range(6):
            left = gap[p, i - 1] if i > 0 else 1.0
            right = gap[p, i] if i < 5 else 1.0
            if min(left, right) > 1e-8:
                d = min(np.abs(Q[p, i] - g["a4eig_Q"][p, i]).max(), np.abs(Q[p, i] + g["a4eig_Q"][p, i]).max())
                assert d < 1e-8 / min(left, right) * 1e-2 + 1e-12, (p, i, d)
        gram = np.einsum("iab,jab->ij", Q[p], Q[p])
        assert np.abs(gram - np.eye(6)).max() < 1e-12
    assert ok.all()   # evolved fabrics are valid states
    # validity on the reference's grid: n00 = 1, n20, n40 (L = 4 states)
    K = len(g["n20"])
    xs = np.zeros((15, K))
    xs[0] = 1.0
    xs[1] = g["n20"].real
    xs[6] = g["n40"].real
    ok = np.zeros(K, np.int32)
    lib.emul_a4eig(4, K, xs, np.zeros((K, 6, 3, 3)), np.zeros((K, 6)), ok)
    assert (ok.astype(bool) == g["isvalid"]).all(), int((ok.astype(bool) != g["isvalid"]).sum())
    assert 0.2 < ok.mean() < 0.8
