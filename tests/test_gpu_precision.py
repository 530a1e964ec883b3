"""Error of each output of the mixed-precision stepping path against the
reference FP64 Fortran: 128 parcels of the L = 12 Eulerian field (LROT + REG
+ DDRX, RK4 dt = 0.5, 200 steps) from tests/golden/field_L12_out.npz, whose
final states, a2, a4, eigenframe, Eij (a2 eigenframe and Cartesian, ice
linear grains, alpha = 0.0125) and elastic Vi (along z, Bennett (1968) ice,
Hill average) were written by oracle/_ref (src/moments.f90:37-55,
src/frames.f90:14-22, src/enhancementfactors.f90:23-69,
src/wavepropagation.f90:121-144).

The states are evolved on the GPU in FP32C (FP32 operator application,
FP64 state with the compensated RK update) and in FP64, and every output is
then computed on the GPU in FP64 from the evolved state.  Errors per parcel:
nlm, a2, a4 relative to the parcel's largest entry; Eij and Vi the largest
componentwise relative error.  Eigenframe Eij is compared where the a2
eigen-gap exceeds 1e-3 (below it the frame is ill-conditioned: a state error
e turns the frame by ~e / gap).  Bars: FP64 1e-10 on everything; FP32C about
3x the measured maximum of each output (the measured table is printed)."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

# FP32C: max relative error per output over the 128 parcels (bar)
BARS_FP32C = {"nlm": 2.5e-7, "a2": 1.5e-7, "a4": 1.5e-7, "Eij_xyz": 1.2e-7, "Eij_eig": 1.2e-7, "Vi": 2.5e-9}


def _outputs(g, precision):
    import torch
    from specfab_amd.batch import FabricBatch
    b = FabricBatch(12, nlm=g["nlm0"], ugrad=g["ugrad"], gam0=g["gam0"])
    b.evolve(int(g["nsteps"]), float(g["dt"]), scheme="rk4", procs=("LROT", "REG", "DDRX"), precision=precision)
    Eg = tuple(g["Eij_grain"])
    E, lam, ei = b.Eij(Eij_grain=Eg, alpha=float(g["alpha"]), n_grain=1)
    Ex = b.Eij(Eij_grain=Eg, alpha=float(g["alpha"]), n_grain=1, frame="xyz")
    V = b.Vi_elastic(float(g["vi_alpha"]), list(g["Lame"]), float(g["rho"]), [0.0], [0.0])
    out = {"nlm": b.nlm(), "a2": b.a2(), "a4": b.a4(), "Eij_eig": E, "Eij_xyz": Ex, "Vi": V[:, 0, :], "lam": lam}
    torch.cuda.synchronize()
    assert not b.status.cpu().numpy().any()
    return {k: v.cpu().numpy() for k, v in out.items()}


def _errors(out, g):
    P = g["nlm"].shape[0]

    def rel_block(a, r):
        a, r = a.reshape(P, -1), r.reshape(P, -1)
        return np.abs(a - r).max(1) / np.abs(r).max(1)

    def rel_comp(a, r):
        return (np.abs(a - r) / np.abs(r)).reshape(a.shape[0], -1).max(1)

    gap = np.minimum(g["lami"][:, 0] - g["lami"][:, 1], g["lami"][:, 1] - g["lami"][:, 2])
    ok = gap > 1e-3
    return {"nlm": rel_block(out["nlm"], g["nlm"]), "a2": rel_block(out["a2"], g["a2"]),
            "a4": rel_block(out["a4"], g["a4"]), "Eij_xyz": rel_comp(out["Eij_xyz"], g["Eij_xyz"]),
            "Eij_eig": rel_comp(out["Eij_eig"][ok], g["Eij_eig"][ok]),
            "Vi": rel_comp(out["Vi"], g["Vi"][:, :, 0])}


def test_fp64_outputs_vs_reference(gpu_ctx):
    g = load_golden("field_L12_out.npz")
    err = _errors(_outputs(g, "fp64"), g)
    print("FP64 max rel err: " + ", ".join("%s %.2e" % (k, v.max()) for k, v in err.items()))
    for k, v in err.items():
        assert v.max() < 1e-10, k


def test_fp32c_error_per_output(gpu_ctx):
    g = load_golden("field_L12_out.npz")
    err = _errors(_outputs(g, "fp32c"), g)
    print("FP32C vs reference FP64, relative error over %d parcels (max / median):" % g["nlm"].shape[0])
    for k, v in err.items():
        print("  %-8s %.2e / %.2e  (bar %.0e)" % (k, v.max(), np.median(v), BARS_FP32C[k]))
    for k, v in err.items():
        assert v.max() < BARS_FP32C[k], k
