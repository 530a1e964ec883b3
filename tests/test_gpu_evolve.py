"""GPU parity of the batched stepping path (libspecfab_gpu through its C-ABI)
against the reference's golden trajectories and the C oracle.

Tolerance: FP64 path, relative max-norm error <= 1e-10 against the reference
FP64 Fortran (tests/golden, oracle/_ref) and the C oracle."""
import numpy as np
import pytest

from conftest import load_golden, relerr

pytestmark = pytest.mark.gpu
TOL = 1e-10


def run(L, nlm0, ugrad, gam0, lam, nsteps, dt, scheme, procs, steps=None, precision="fp64", iota=1.0, zeta=0.0):
    import torch
    from specfab_amd.batch import FabricBatch
    b = FabricBatch(L, nlm=nlm0, ugrad=ugrad, gam0=gam0, lam=lam, steps=steps)
    b.evolve(nsteps, dt, scheme=scheme, procs=procs, precision=precision, iota=iota, zeta=zeta)
    out = b.nlm().cpu().numpy()
    st = b.status.cpu().numpy()
    torch.cuda.synchronize()
    return out, st


def test_state_roundtrip(gpu_ctx):
    from specfab_amd import workloads as W
    from specfab_amd.batch import FabricBatch
    rng = np.random.default_rng(0)
    for L in (2, 4, 8, 12, 20):
        nlm = W.perturbed_isotropic(rng, 33, L, sigma=0.1)
        b = FabricBatch(L, nlm=nlm, ugrad=np.zeros((33, 3, 3)))
        assert np.array_equal(b.nlm().cpu().numpy(), nlm)


def test_c1_euler_trajectory(gpu_ctx):
    g = load_golden("c1_traj.npz")
    x = g["nlm0"]
    for t in range(1, g["traj"].shape[0]):
        x, st = run(8, x, g["ugrad"], 0.0, 0.0, int(g["every"]), float(g["dt"]), "euler", ("LROT", "REG"))
        assert relerr(x[0], g["traj"][t]) < TOL, t
        assert st[0] == 0


def test_field_rk4_L12(gpu_ctx):
    g = load_golden("field_L12.npz")
    out, st = run(12, g["nlm0"], g["ugrad"], g["gam0"], 0.0, int(g["nsteps"]), float(g["dt"]), "rk4",
                  ("LROT", "REG", "DDRX"))
    for p in range(out.shape[0]):
        assert relerr(out[p], g["nlm"][p]) < TOL, p
    assert not st.any()


def test_uneven_per_L(gpu_ctx):
    g = load_golden("uneven.npz")
    for L in (4, 8, 12, 20):
        out, _ = run(L, g["nlm0_L%d" % L], g["ugrad_L%d" % L], g["gam0_L%d" % L], g["lam_L%d" % L], 0, 0.5, "rk4",
                     ("LROT", "REG", "DDRX", "CDRX"), steps=g["steps_L%d" % L])
        for k in range(out.shape[0]):
            assert relerr(out[k], g["nlm_L%d" % L][k]) < TOL, (L, k)


def test_euler_ddrx(gpu_ctx):
    g = load_golden("euler_ddrx_L8.npz")
    out, _ = run(8, g["nlm0"], g["ugrad"], g["gam0"], 0.0, int(g["nsteps"]), float(g["dt"]), "euler",
                 ("LROT", "REG", "DDRX"))
    assert relerr(out, g["nlm"]) < TOL


@pytest.mark.parametrize("L,procs", [(12, ("LROT", "REG", "DDRX")), (8, ("LROT", "DDRX", "CDRX")),
                                     (4, ("LROT", "REG", "DDRX", "CDRX")), (20, ("LROT", "REG"))])
def test_random_batch_vs_oracle(gpu_ctx, oracle, L, procs):
    from specfab_amd import workloads as W
    rng = np.random.default_rng(L + 100)
    P = 300
    nlm0 = W.perturbed_isotropic(rng, P, L, sigma=1e-2)
    ug = W.traceless_gradients(rng, P)
    g0 = 10.0 ** rng.uniform(-4, -2, P)
    lam = 10.0 ** rng.uniform(-4, -3, P)
    steps = 12
    out, st = run(L, nlm0, ug, g0, lam, steps, 0.5, "rk4", procs)
    ref = oracle.evolve(L, nlm0, ug, g0, lam, steps, 0.5, "rk4", procs)
    assert relerr(out, ref) < TOL
    assert not st.any()


def test_zeta_and_iota(gpu_ctx, oracle):
    from specfab_amd import workloads as W
    rng = np.random.default_rng(5)
    nlm0 = W.perturbed_isotropic(rng, 20, 8, sigma=1e-2)
    ug = W.traceless_gradients(rng, 20)
    out, _ = run(8, nlm0, ug, 0.0, 0.0, 30, 1.0, "rk4", ("LROT", "REG"), iota=0.7, zeta=0.3)
    ref = oracle.evolve(8, nlm0, ug, 0.0, 0.0, 30, 1.0, "rk4", ("LROT", "REG"), iota=0.7, zeta=0.3)
    assert relerr(out, ref) < TOL


def test_edge_cases(gpu_ctx):
    from specfab_amd import _lib
    from specfab_amd.batch import FabricBatch
    # empty batch is a no-op
    b = FabricBatch(8, P=0, ugrad=np.zeros((0, 3, 3)))
    b.evolve(5, 1.0)
    # zero strain rate: zetanorm = 0/0 makes the reference's M_LROT NaN
    # (src/dynamics.f90:73); the status flags it instead of stopping
    b = FabricBatch(8, P=3, ugrad=np.zeros((3, 3, 3)))
    b.evolve(2, 1.0, procs=("LROT",))
    assert (b.status.cpu().numpy() & 1).all()
    # REG is calibrated only for L = 4..20 (src/include/regcalib.f90)
    b = FabricBatch(2, P=2, ugrad=np.ones((2, 3, 3)) * 1e-3)
    with pytest.raises(_lib.SFGError):
        b.evolve(1, 1.0, procs=("LROT", "REG"))


def test_fp32c_field_error_vs_fp64(gpu_ctx, oracle):
    """FP32 stepping with the FP64-accumulated state: error against the FP64
    oracle on the L = 12 field (LROT + REG + DDRX, RK4, 200 steps)."""
    g = load_golden("field_L12.npz")
    out, st = run(12, g["nlm0"], g["ugrad"], g["gam0"], 0.0, int(g["nsteps"]), float(g["dt"]), "rk4",
                  ("LROT", "REG", "DDRX"), precision="fp32c")
    err = max(relerr(out[p], g["nlm"][p]) for p in range(out.shape[0]))
    print("fp32c max rel err vs FP64 reference: %.3e" % err)
    assert err < 1e-5
    assert not st.any()


@pytest.mark.parametrize("precision,tol", [("fp64", TOL), ("fp32c", 1e-5)])
def test_full_size_field_sample(gpu_ctx, oracle, precision, tol):
    """The headline field at full size (2^20 parcels, L = 12: a 763 MB state,
    LROT + REG + DDRX, RK4 dt = 0.5, 20 steps) on the autotuned stepper.
    Parcels are independent, so a sample of 64 of them (first and last,
    wavefront and 2^16 boundaries, random others) stepped alone must be bit
    for bit the same as inside the full batch (a check of the SoA indexing
    over the whole state), and matches the oracle; over all 2^20 parcels n00
    stays put (the LROT and DDRX source rows vanish up to the float32
    rounding of the coupling tables) and no status bit is set."""
    import math
    import torch
    from specfab_amd import _lib
    from specfab_amd.batch import FabricBatch
    dev = torch.device("cuda:0")
    P, L, nsteps, procs = 1 << 20, 12, 20, ("LROT", "REG", "DDRX")
    n = (L + 1) * (L + 2) // 2
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    G = torch.randn((P, 3, 3), generator=g, device=dev, dtype=torch.float64) * 1e-3
    G -= (G.diagonal(dim1=1, dim2=2).sum(-1) / 3)[:, None, None] * torch.eye(3, device=dev, dtype=torch.float64)
    gam0 = 10.0 ** (torch.rand(P, generator=g, device=dev, dtype=torch.float64) * 2 - 4)
    x0 = torch.randn((n, P), generator=g, device=dev, dtype=torch.float64) * 1e-3
    x0[0] = 1 / math.sqrt(4 * math.pi)

    def batch(x, ug, g0):
        b = FabricBatch.__new__(FabricBatch)
        b.ctx = _lib.default_context()
        b.L, b.n, b.dev, b.P = L, n, dev, x.shape[1]
        b.x, b.ugrad, b.tau, b.gam0, b.lam, b.steps = x.contiguous(), ug.contiguous(), None, g0.contiguous(), None, None
        b.status = torch.zeros(b.P, dtype=torch.int32, device=dev)
        return b

    fb = batch(x0.clone(), G, gam0)
    fb.autotune(precision, P)
    fb.evolve(nsteps, 0.5, scheme="rk4", procs=procs, precision=precision)
    rng = np.random.default_rng(11)
    edges = [0, 1, 63, 64, 65535, 65536, P // 2, P - 65, P - 64, P - 1]
    sample = np.unique(np.concatenate([edges, rng.choice(P, 64 - len(edges), replace=False)]))
    si = torch.from_numpy(sample).to(dev)
    alone = batch(x0[:, si], G[si], gam0[si])
    nlm0 = alone.nlm().cpu().numpy()
    alone.evolve(nsteps, 0.5, scheme="rk4", procs=procs, precision=precision)
    assert torch.equal(alone.x, fb.x[:, si])
    ref = oracle.evolve(L, nlm0, G[si].cpu().numpy(), gam0[si].cpu().numpy(), 0.0, nsteps, 0.5, "rk4", procs)
    err = relerr(alone.nlm().cpu().numpy(), ref)
    drift = ((fb.x[0] - x0[0]).abs() / x0[0]).max().item()
    print("full field %s: sample vs oracle %.2e, max n00 drift %.2e" % (precision, err, drift))
    assert err < tol
    assert drift < 1e-6
    assert not fb.status.any().item()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32c", 1e-7)])
def test_large_batch_in_slices(gpu_ctx, oracle, precision, tol):
    """2^22 parcels at L = 12 (a 3.05 GB state, more than the 2^31 bytes the
    steppers address with 32-bit buffer offsets) on the generated steppers:
    the call steps the batch in slices (2,949,632 parcels: each slice's state
    staged in the workspace and copied back).  64 sampled parcels (first,
    last, both sides of the slice boundary, random others) stepped alone are
    bit for bit the same as inside the batch, and match the oracle; no status
    bit is set."""
    import math
    import torch
    from specfab_amd import _lib
    from specfab_amd.batch import FabricBatch
    dev = torch.device("cuda:0")
    P, L, nsteps, procs = 1 << 22, 12, 3, ("LROT", "REG", "DDRX")
    n = (L + 1) * (L + 2) // 2
    slice_ = ((2 ** 31 - 1) // (8 * n)) // 256 * 256
    assert n * P * 8 >= 2 ** 31 and slice_ < P
    print("large batch %s: P=%d, slice %d" % (precision, P, slice_), flush=True)
    g = torch.Generator(device=dev)
    g.manual_seed(23)
    G = torch.randn((P, 3, 3), generator=g, device=dev, dtype=torch.float64) * 1e-3
    G -= (G.diagonal(dim1=1, dim2=2).sum(-1) / 3)[:, None, None] * torch.eye(3, device=dev, dtype=torch.float64)
    gam0 = 10.0 ** (torch.rand(P, generator=g, device=dev, dtype=torch.float64) * 2 - 4)
    x0 = torch.randn((n, P), generator=g, device=dev, dtype=torch.float64) * 1e-3
    x0[0] = 1 / math.sqrt(4 * math.pi)

    def batch(x, ug, g0):
        b = FabricBatch.__new__(FabricBatch)
        b.ctx = _lib.default_context()
        b.L, b.n, b.dev, b.P = L, n, dev, x.shape[1]
        b.x, b.ugrad, b.tau, b.gam0, b.lam, b.steps = x.contiguous(), ug.contiguous(), None, g0.contiguous(), None, None
        b.status = torch.zeros(b.P, dtype=torch.int32, device=dev)
        return b

    fb = batch(x0.clone(), G, gam0)
    fb.autotune(precision, 1 << 16, steps=nsteps)
    fb.evolve(nsteps, 0.5, scheme="rk4", procs=procs, precision=precision)
    rng = np.random.default_rng(12)
    edges = [0, 1, 63, slice_ - 2, slice_ - 1, slice_, slice_ + 1, P - 64, P - 1]
    sample = np.unique(np.concatenate([edges, rng.choice(P, 64 - len(edges), replace=False)]))
    si = torch.from_numpy(sample).to(dev)
    alone = batch(x0[:, si], G[si], gam0[si])
    nlm0 = alone.nlm().cpu().numpy()
    alone.evolve(nsteps, 0.5, scheme="rk4", procs=procs, precision=precision)
    assert torch.equal(alone.x, fb.x[:, si])
    ref = oracle.evolve(L, nlm0, G[si].cpu().numpy(), gam0[si].cpu().numpy(), 0.0, nsteps, 0.5, "rk4", procs)
    err = relerr(alone.nlm().cpu().numpy(), ref)
    print("large batch %s: sample vs oracle %.2e" % (precision, err))
    assert err < tol
    assert not fb.status.any().item()
    del fb, x0, G, gam0
    torch.cuda.empty_cache()


def test_fast_vs_generic_stepper(gpu_ctx):
    """The generated register-resident stepper and the table-driven generic
    stepper agree on the GPU (same inputs, FP64)."""
    from specfab_amd import _lib
    from specfab_amd import workloads as W
    from specfab_amd.batch import FabricBatch
    rng = np.random.default_rng(77)
    for L in (4, 8, 12):
        P = 130
        nlm0 = W.perturbed_isotropic(rng, P, L, sigma=1e-2)
        ug = W.traceless_gradients(rng, P)
        g0 = 10.0 ** rng.uniform(-4, -2, P)
        outs = []
        for stepper in ("fast", "generic"):
            ctx = _lib.Context(0)
            if stepper == "generic":
                import os
                os.environ["SFG_STEPPER"] = "generic"
                ctx = _lib.Context(0)
                del os.environ["SFG_STEPPER"]
            b = FabricBatch(L, nlm=nlm0, ugrad=ug, gam0=g0, ctx=ctx)
            b.evolve(7, 0.5, scheme="rk4", procs=("LROT", "REG", "DDRX"))
            outs.append(b.nlm().cpu().numpy())
        print("L=%d fast vs generic %.3e" % (L, relerr(outs[0], outs[1])))
        assert relerr(outs[0], outs[1]) < 1e-12, L


@pytest.mark.parametrize("L", [8, 12, 14, 20])
def test_every_stepper_variant(gpu_ctx, oracle, L):
    """Every generated stepper variant (FP64 and FP32 operator) against the
    oracle, and the autotuner's choice is one of them; L = 14 and 20 cover
    the window steppers, including those that split a parcel group's row
    chunks over several waves with workgroup barriers (P = 200: three full
    64-parcel groups and a partial one); RK4 and explicit Euler (whose
    window steppers write the new state to a scratch buffer and copy it back
    per wave)."""
    import os
    from specfab_amd import _lib
    from specfab_amd import workloads as W
    from specfab_amd.batch import FabricBatch
    rng = np.random.default_rng(300 + L)
    P = 200
    nlm0 = W.perturbed_isotropic(rng, P, L, sigma=1e-2)
    ug = W.traceless_gradients(rng, P)
    g0 = 10.0 ** rng.uniform(-4, -2, P)
    procs = ("LROT", "REG", "DDRX")
    ref = {sc: oracle.evolve(L, nlm0, ug, g0, 0.0, 5, 0.5, sc, procs) for sc in ("rk4", "euler")}
    for prec, tol in (("fp64", 1e-12), ("fp32c", 2e-6)):
        b0 = FabricBatch(L, nlm=nlm0, ugrad=ug, gam0=g0)
        best, times = b0.autotune(prec, P=4096)
        print("L=%d %s autotune: variant %d, ms %s" % (L, prec, best, ["%.3f" % t for t in times]))
        assert 0 <= best < len(times)
        for v in range(len(times)):
            os.environ["SFG_VARIANT"] = str(v)
            try:
                ctx = _lib.Context(0)
            finally:
                del os.environ["SFG_VARIANT"]
            for sc in ("rk4", "euler"):
                b = FabricBatch(L, nlm=nlm0, ugrad=ug, gam0=g0, ctx=ctx)
                b.evolve(5, 0.5, scheme=sc, procs=procs, precision=prec)
                err = relerr(b.nlm().cpu().numpy(), ref[sc])
                assert err < tol, (prec, v, sc, err)
            ctx.close()


@pytest.mark.parametrize("L", [12, 20])
@pytest.mark.parametrize("precision,tol", [("fp64", TOL), ("fp32c", 2e-6)])
def test_uneven_plan_every_variant(gpu_ctx, precision, tol, L):
    """The L = 12 and L = 20 parcels of the uneven fixture, each repeated 200
    times and shuffled (600 parcels: ten parcel groups of 64, per-parcel step
    counts, so the lanes of one wave stop at different steps), through a
    graph plan with every stepper variant of that truncation forced in turn,
    including those that split a parcel group's rows over 2, 4 or 8 waves
    with workgroup barriers in the step loop (L = 12 resident steppers, L =
    20 window steppers): final states against the reference trajectories,
    every step counter at zero."""
    import os
    from specfab_amd import _lib
    from specfab_amd.batch import FabricBatch, UnevenBatch
    g = load_golden("uneven.npz")
    rep = np.random.default_rng(7).permutation(np.repeat(np.arange(len(g["steps_L%d" % L])), 200))
    nlm0, ug, g0, lam, steps, ref = (g["%s_L%d" % (k, L)][rep] for k in ("nlm0", "ugrad", "gam0", "lam", "steps", "nlm"))
    _, times = FabricBatch(L, nlm=nlm0, ugrad=ug, gam0=g0).autotune(precision, P=256)
    nv = len(times)
    assert nv >= 3
    for v in range(nv):
        os.environ["SFG_VARIANT"] = str(v)
        try:
            ctx = _lib.Context(0)
        finally:
            del os.environ["SFG_VARIANT"]
        ub = UnevenBatch([L] * len(steps), list(nlm0), np.array(ug), np.array(g0), np.array(lam), np.array(steps),
                         ctx=ctx)
        ub.make_plan(0.5, "rk4", ("LROT", "REG", "DDRX", "CDRX"), precision=precision, chunk=64).run()
        out = ub.nlm()
        err = max(relerr(out[k], ref[k]) for k in range(len(steps)))
        print("L=%d %s variant %d: max rel err %.2e over %d parcels, steps %d..%d"
              % (L, precision, v, err, len(steps), steps.min(), steps.max()))
        assert err < tol, (precision, v, err)
        assert (ub.remaining() == 0).all(), (precision, v)
        ub.close()
        ctx.close()


@pytest.mark.parametrize("precision,tol", [("fp64", TOL), ("fp32c", 2e-6)])
def test_uneven_plan_vs_reference(gpu_ctx, precision, tol):
    """Uneven batch (L in {4, 8, 12, 20} mixed, 50..500 steps per parcel,
    LROT+REG+DDRX+CDRX, RK4) through one graph launch, in a shuffled parcel
    order, against the reference trajectories."""
    from specfab_amd.batch import UnevenBatch
    g = load_golden("uneven.npz")
    Ls_all = (4, 8, 12, 20)
    Ls, nlm0, ug, g0, lam, steps, ref = [], [], [], [], [], [], []
    for L in Ls_all:
        for k in range(g["nlm0_L%d" % L].shape[0]):
            Ls.append(L)
            nlm0.append(g["nlm0_L%d" % L][k])
            ug.append(g["ugrad_L%d" % L][k])
            g0.append(g["gam0_L%d" % L][k])
            lam.append(g["lam_L%d" % L][k])
            steps.append(g["steps_L%d" % L][k])
            ref.append(g["nlm_L%d" % L][k])
    perm = np.random.default_rng(0).permutation(len(Ls))
    sel = lambda a: [a[i] for i in perm]
    ub = UnevenBatch(sel(Ls), sel(nlm0), np.array(sel(ug)), np.array(sel(g0)), np.array(sel(lam)),
                     np.array(sel(steps)))
    ub.make_plan(0.5, "rk4", ("LROT", "REG", "DDRX", "CDRX"), precision=precision, chunk=64).run()
    out = ub.nlm()
    for k, i in enumerate(perm):
        assert relerr(out[k], ref[i]) < tol, (Ls[i], precision, relerr(out[k], ref[i]))
    assert (ub.remaining() == 0).all()
    assert not ub.status().any()
    # a second launch does nothing (no steps remain)
    ub.run()
    out2 = ub.nlm()
    assert all(np.array_equal(a, b) for a, b in zip(out, out2))
