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
    lib.emul_a4.argtypes = [c.c_int, c.c_int64, d, d]
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
