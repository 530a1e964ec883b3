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


def test_fortran_c1_parcel():
    if not os.path.exists(EXE):
        pytest.skip("Fortran demo not built (amdflang missing at build time)")
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = [l.split() for l in out.stdout.splitlines() if l.startswith("nlm ")]
    host = np.array([float(r[2]) + 1j * float(r[3]) for r in rows])
    batched = np.array([float(r[4]) + 1j * float(r[5]) for r in rows])
    ref = load_golden("c1_traj.npz")["traj"][-1]
    assert relerr(host, ref) < 1e-10
    assert relerr(batched, ref) < 1e-10
