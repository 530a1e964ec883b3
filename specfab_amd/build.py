"""Build libspecfab_gpu.so in-tree (hipcc, gfx950) and its generated sources.

    python -m specfab_amd.build      (or __graft_entry__.build())

Steps: (1) the coupling tables (specfab_amd/tables/make_gaunt.py) are
generated into csrc/generated/ unless present; (2) the register-resident
steppers are generated (specfab_amd/tables/gen_fast.py); (3) all HIP/C++
sources are compiled into specfab_amd/libspecfab_gpu.so.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
GEN = os.path.join(CSRC, "generated")
LIB = os.path.join(HERE, "libspecfab_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SFG_ARCH", "gfx950")

SOURCES = ["sfg_basis.cpp", "sfg_wave_host.cpp", "sfg_api.cpp", "sfg_kernels.hip", "sfg_homog.hip", "sfg_wave.hip", "sfg_mlp_host.cpp", "sfg_mlp.hip", "sfg_surrogate.hip", "sfg_init.hip", "sfg_reduced.hip", "sfg_rotate.hip", "sfg_noalloc.hip", "sfg_discrete.hip"]
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-Wno-unused-result"]


def run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, **kw)


def gen_tables(force=False):
    os.makedirs(GEN, exist_ok=True)
    inc = os.path.join(GEN, "gaunt_L20.inc")
    if force or not os.path.exists(inc):
        run([sys.executable, os.path.join(HERE, "tables", "make_gaunt.py"), inc])
    run([sys.executable, os.path.join(HERE, "tables", "make_a6_map.py"), os.path.join(GEN, "sfg_a6_map.h")])
    run([sys.executable, os.path.join(HERE, "tables", "make_moment_maps.py"), os.path.join(GEN, "sfg_moments68.h")])
    # the l <= 4 a2 / a4 maps as compile-time tables (host program over the map code)
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    gen = os.path.join(objdir, "gen_l4maps")
    run([os.environ.get("CXX_HOST", "g++"), "-std=c++17", "-O1", "-I", CSRC,
         os.path.join(HERE, "tables", "gen_l4maps.cpp"), os.path.join(CSRC, "sfg_wave_host.cpp"),
         os.path.join(CSRC, "sfg_basis.cpp"), "-o", gen])
    run([gen, os.path.join(GEN, "sfg_l4maps.h")])
    gen_fast = os.path.join(HERE, "tables", "gen_fast.py")
    if os.path.exists(gen_fast):
        run([sys.executable, gen_fast, GEN])


def compile_lib(jobs=8):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = list(SOURCES)
    for f in sorted(os.listdir(GEN)):
        if f.endswith(".hip"):
            srcs.append(os.path.join("generated", f))
    procs, objs = [], []
    for src in srcs:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        extra = []
        if src.startswith("generated" + os.sep + "sfg_fast_L"):
            # the steppers are one long basic block per row chunk: keep the
            # literal constants next to their uses (no loop-invariant hoisting)
            extra = ["-mllvm", "-disable-machine-licm"]
        cmd = [HIPCC] + HIPFLAGS + extra + ["-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            import time
            time.sleep(0.2)
    bad = [s for s, p in procs if p.wait() != 0]
    if bad:
        raise RuntimeError("compile failed: %s" % bad)
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", "-fPIC", "-Wl,--no-undefined", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def compile_fortran():
    """Fortran face (ISO_C_BINDING module specfab_gpu) and its test program,
    built with amdflang against libspecfab_gpu.so."""
    fc = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")
    if not os.path.exists(fc):
        print("amdflang not found; skipping the Fortran module")
        return None
    out = os.path.join(HERE, "fortran", "_build")
    os.makedirs(out, exist_ok=True)
    rocm_lib = "/opt/rocm/lib"
    exe = os.path.join(out, "c1_parcel")
    run([fc, "-O2", "-J", out, "-c", os.path.join(HERE, "fortran", "specfab_gpu.f90"), "-o",
         os.path.join(out, "specfab_gpu.o")])
    for prog in ("c1_parcel", "homog_faces", "discrete_faces", "ortho_faces", "faces"):
        run([fc, "-O2", "-I", out, os.path.join(os.path.dirname(HERE), "tests", "fortran", prog + ".f90"),
             os.path.join(out, "specfab_gpu.o"), "-L", HERE, "-lspecfab_gpu", "-L", rocm_lib, "-lamdhip64",
             "-Wl,-rpath," + HERE, "-Wl,-rpath," + rocm_lib, "-o", os.path.join(out, prog)])
    return exe


def main():
    gen_tables()
    print("built", compile_lib())
    compile_fortran()


if __name__ == "__main__":
    main()
