#!/usr/bin/env python3
"""Benchmark: batched fabric evolution of an Eulerian CPO field on one MI355X.

Workload (the headline configuration): P = 2^20 parcels at L = 12 (91 complex
coefficients), per-parcel traceless velocity gradients ~ N(0, 1e-3),
DDRX rate Gamma0 ~ LogUniform(1e-4, 1e-2), lattice rotation + regularization
+ DDRX, classic RK4 with dt = 0.5.  One "step" is one RK4 step of the whole
field; the full configuration runs 200 of them, this script times --steps.
Inputs are synthetic, generated on the GPU from a fixed seed, and resident
in HBM before the timed region starts.

Prints one JSON line (metric parcel-steps/s) with a roofline object for the
stepping launch (FP64 vector-ALU bound) and the reference CPU path timed on
the host's cores on a bounded sample.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "parcel-steps/s at L=12 + Eij evals/s, plain vs learned at ≤1e-3 rel err"
# MI355X published dense peaks: FP64 (vector and matrix alike) 78.6 TFLOP/s,
# FP32 157.3 TFLOP/s.
PEAK_TFLOPS = {"fp64": 78.6, "fp32c": 157.3}


def flops_per_parcel_step(L=12, stages=4, ddrx=True):
    """Algorithmic flops of one RK4 parcel-step in the real basis:
    stages * (2 * T_r + 4 n + 300 [<D>]) + 10 n (RK combination).
    T_r = coupling nonzeros of LROT + DDRX: 8,581 at L = 12."""
    T_r = {4: 679, 8: 3467, 12: 8581, 20: 25773}[L]
    n = (L + 1) * (L + 2) // 2
    return stages * (2 * T_r + 4 * n + (300 if ddrx else 0)) + 10 * n


def make_field(P, L, seed, device):
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    n = (L + 1) * (L + 2) // 2
    G = torch.randn((P, 3, 3), generator=g, device=device, dtype=torch.float64) * 1e-3
    tr = G.diagonal(dim1=1, dim2=2).sum(-1)
    G -= (tr / 3)[:, None, None] * torch.eye(3, device=device, dtype=torch.float64)
    gam0 = 10.0 ** (torch.rand(P, generator=g, device=device, dtype=torch.float64) * 2 - 4)
    x = torch.randn((n, P), generator=g, device=device, dtype=torch.float64) * 1e-3
    x[0] = 1 / math.sqrt(4 * math.pi)
    return G.contiguous(), gam0.contiguous(), x.contiguous()


def cpu_worker(args):
    """Reference Fortran (or C oracle) on one host core: RK4 of a slice."""
    kind, P, steps, seed = args
    import numpy as np
    sys.path.insert(0, ROOT)
    from specfab_amd import workloads as W
    f = W.eulerian_field(P, L=12, seed=seed, nsteps=steps)
    if kind == "reference":
        from oracle.ref import RefLib
        r = RefLib()
        r.init(12)
        t0 = time.perf_counter()
        r.evolve(f["nlm"], f["ugrad"], f["gam0"], f["lam"], steps, f["dt"], scheme="rk4", procs=f["procs"])
    else:
        from oracle.oracle import Oracle
        o = Oracle()
        t0 = time.perf_counter()
        o.evolve(12, f["nlm"], f["ugrad"], f["gam0"], f["lam"], steps, f["dt"], "rk4", f["procs"])
    return time.perf_counter() - t0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(target_s=10.0):
    """The reference Fortran (oracle/_ref; the C restatement where it is not
    built) on the host: one process per core (the reference keeps global
    state and is not thread-safe), at most 16 processes: the CPU share that
    goes with one GPU of the box, whatever the affinity set shows."""
    import multiprocessing as mp
    from oracle import ref as refmod
    kind = "reference" if refmod.available() else "port"
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    ncores = max(1, min(affinity, 16))
    # calibrate on one core: 16 parcels x 4 steps (stepping time only)
    t1 = cpu_worker((kind, 16, 4, 99)) / 64
    steps = 8
    per_proc = max(1, int(target_s / (t1 * steps)))   # each core runs ~target_s of stepping
    ctx = mp.get_context("spawn")
    with ctx.Pool(ncores) as pool:
        times = pool.map(cpu_worker, [(kind, per_proc, steps, 1000 + i) for i in range(ncores)])
    total = ncores * per_proc * steps
    slowest = max(times)
    return {"value": total / slowest, "unit": "parcel-steps/s", "cores": ncores, "kind": kind,
            "affinity_cores": affinity, "cpu_model": cpu_model(), "single_core_value": 1.0 / t1,
            "sample": "%d processes (one per core) x %d parcels x %d RK4 steps at L=12 (LROT+REG+DDRX), "
                      "stepping time of the slowest process %.1f s; single-core rate %.1f parcel-steps/s"
                      % (ncores, per_proc, steps, slowest, 1.0 / t1)}


def bench_uneven(a):
    """Uneven batch: 2^18 parcels, L ~ U{4, 8, 12, 20}, steps ~ U[50, 500], traceless
    N(0, 1e-3) gradients, Gamma0 ~ 10^U(-4,-2), Lambda ~ 10^U(-4,-3),
    LROT+REG+DDRX+CDRX, RK4 dt = 0.5, whole time loop as one graph launch."""
    import numpy as np
    import torch
    from specfab_amd import batch as B
    P = 1 << 18
    rng = np.random.default_rng(4)
    Ls = rng.choice([4, 8, 12, 20], P)
    steps = rng.integers(50, 501, P)
    G = rng.normal(0, 1e-3, (P, 3, 3))
    G -= np.trace(G, axis1=1, axis2=2)[:, None, None] / 3 * np.eye(3)
    g0 = 10.0 ** rng.uniform(-4, -2, P)
    lam = 10.0 ** rng.uniform(-4, -3, P)
    nlm = []
    for L in Ls:
        n = (L + 1) * (L + 2) // 2
        x = np.zeros(n, complex)
        x[0] = 1 / math.sqrt(4 * math.pi)
        nlm.append(x)
    ub = B.UnevenBatch(Ls, nlm, G, g0, lam, steps)
    for L, idx, fb, rem in ub.buckets:
        fb.autotune(a.precision, steps=10)   # the plan launches 50 steps at a time
    ub.make_plan(0.5, "rk4", ("LROT", "REG", "DDRX", "CDRX"), precision=a.precision, chunk=50)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ub.run()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    work = int(steps.sum())
    flops = sum(flops_per_parcel_step(L) * int(steps[Ls == L].sum()) for L in (4, 8, 12, 20))
    per_L = {int(L): int(steps[Ls == L].sum()) for L in (4, 8, 12, 20)}
    out = {"metric": "parcel-steps/s, uneven batch (mixed L and step counts) in one graph launch",
           "value": work / wall, "unit": "parcel-steps/s", "n_gpus": 1, "steps": 1, "warmup": 0,
           "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if a.precision == "fp64" else "f32+compensated", "data": "synthetic (seeded)",
           "config": {"workload": "uneven batch: 2^18 parcels, L ~ U{4,8,12,20}, steps ~ U[50,500], "
                                  "LROT+REG+DDRX+CDRX, RK4 dt=0.5, one graph launch", "parcels": P,
                      "parcel_steps": work, "parcel_steps_by_L": per_L},
           "achieved_tflops": flops / wall / 1e12,
           "nonfinite_parcels": int((ub.status() != 0).sum())}
    print(json.dumps(out), flush=True)


def bench_surrogate(a):
    """In-loop homogenization of the 2^20-parcel L = 12 field after 200 RK4
    steps: plain FP64 Eij (Cartesian) + Vi (z) vs the trained network with
    envelope fallback and a 1 % audit."""
    import torch
    from specfab_amd import surrogate as S
    P = a.parcels
    X, T, fb = S.field_pairs(P, steps=200, seed=11)
    del X, T
    net = S.Surrogate().load()
    hp = S.ice_params()
    E0, E1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def plain_homog():
        fb.Eij((hp.Eij_grain[0], hp.Eij_grain[1]), hp.alpha, hp.n_grain, frame="xyz")
        fb.Vi_elastic(hp.vi_alpha, list(hp.Lame), hp.rho, [0.0], [0.0])

    plain_homog()
    torch.cuda.synchronize()
    E0.record()
    for _ in range(3):
        plain_homog()
    E1.record()
    torch.cuda.synchronize()
    t_plain = E0.elapsed_time(E1) / 3
    net.homogenize(fb, hp, audit_frac=0.01)
    torch.cuda.synchronize()
    E0.record()
    for _ in range(3):
        E, V, st = net.homogenize(fb, hp, audit_frac=0.01)
    E1.record()
    torch.cuda.synchronize()
    t_sur_stats = E0.elapsed_time(E1) / 3
    # the same without fetching the audit report (stream-ordered, no host sync)
    E0.record()
    for _ in range(3):
        net.homogenize(fb, hp, audit_frac=0.01, stats=False)
    E1.record()
    torch.cuda.synchronize()
    t_sur = E0.elapsed_time(E1) / 3
    _, Ep, Vp = S.plain_targets(fb, hp)
    err = ((E - Ep).abs() / Ep.abs()).max(1).values
    # in-loop use: one RK4 step of the field followed by its homogenization,
    # plain FP64 (stepping + Eij + Vi) vs FP32 stepping + the network
    procs = ("LROT", "REG", "DDRX")
    fb.autotune("fp64", P, steps=1)    # the loop below makes one-step calls
    fb.autotune("fp32c", P, steps=1)

    def plain_step():
        fb.evolve(1, 0.5, scheme="rk4", procs=procs, precision="fp64")
        plain_homog()

    # a fresh 1 % audit sample every step; its statistics stay on the device
    # (no per-step synchronisation)
    seeds = iter(range(1000, 10 ** 6))

    def mixed_step():
        fb.evolve(1, 0.5, scheme="rk4", procs=procs, precision="fp32c")
        net.homogenize(fb, hp, audit_frac=0.01, seed=next(seeds), stats=False)

    loop_ms = {}
    for tag, fn in (("plain_fp64", plain_step), ("fp32c_surrogate", mixed_step)):
        fn()
        torch.cuda.synchronize()
        E0.record()
        for _ in range(5):
            fn()
        E1.record()
        torch.cuda.synchronize()
        loop_ms[tag] = E0.elapsed_time(E1) / 5
    # homogenization every 10 steps (an output cadence): 10 steps in one call,
    # then Eij and Vi; steppers re-tuned for 10-step launches; ms per step
    fb.autotune("fp64", P, steps=10)
    fb.autotune("fp32c", P, steps=10)

    def plain_10():
        fb.evolve(10, 0.5, scheme="rk4", procs=procs, precision="fp64")
        plain_homog()

    def mixed_10():
        fb.evolve(10, 0.5, scheme="rk4", procs=procs, precision="fp32c")
        net.homogenize(fb, hp, audit_frac=0.01, seed=next(seeds), stats=False)

    loop10_ms = {}
    for tag, fn in (("plain_fp64", plain_10), ("fp32c_surrogate", mixed_10)):
        fn()
        torch.cuda.synchronize()
        E0.record()
        for _ in range(3):
            fn()
        E1.record()
        torch.cuda.synchronize()
        loop10_ms[tag] = E0.elapsed_time(E1) / 30
    out = {"metric": "Eij + Vi evaluations/s on the L=12 field, plain FP64 vs learned (bf16 MFMA)",
           "value": P / (t_sur * 1e-3), "unit": "parcel-evals/s", "n_gpus": 1, "higher_is_better": True,
           "plain_evals_per_s": P / (t_plain * 1e-3), "speedup_vs_plain": t_plain / t_sur,
           "ms_plain": t_plain, "ms_surrogate": t_sur, "ms_surrogate_with_report": t_sur_stats,
           "dtype": "bf16 (fp32 accumulate)",
           "audit": st, "Eij_rel_err_vs_plain": {"mean": err.mean().item(), "p99": err.quantile(0.99).item(),
                                                   "max": err.max().item()},
           "in_loop_ms_per_step": loop_ms,
           "in_loop_speedup": loop_ms["plain_fp64"] / loop_ms["fp32c_surrogate"],
           "every_10_steps_ms_per_step": loop10_ms,
           "every_10_steps_speedup": loop10_ms["plain_fp64"] / loop10_ms["fp32c_surrogate"],
           "config": {"workload": "Eulerian CPO field, %d parcels, L=12, after 200 RK4 steps; ice linear "
                                  "Sachs/Taylor Eij (Cartesian) + Hill Vi along z" % P, "parcels": P}}
    print(json.dumps(out), flush=True)


def measured_traffic(kernel, P, steps):
    """HBM bytes of one launch of `kernel` on P parcels for `steps` steps, from
    the committed table profiles/evolve_traffic.json: rocprofv3 FETCH_SIZE and
    WRITE_SIZE per dispatch of the bench command, scaled by a plain device
    copy (bytes per KiB counted), fitted per parcel as a launch prologue plus
    a per-step figure from two launch lengths of the same run (so any
    --steps is covered); an entry with a single launch length is used only
    for that length; a resident stepper (state read once per launch, written
    once at its end) has one figure per launch, whatever the step count.
    None for a kernel not measured."""
    path = os.path.join(ROOT, "profiles", "evolve_traffic.json")
    try:
        with open(path) as f:
            tab = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    ent = tab.get(kernel)
    if not ent:
        return None
    if "bytes_per_parcel_launch" in ent:
        return P * ent["bytes_per_parcel_launch"]
    if "prologue_bytes_per_parcel" in ent:
        return P * (ent["prologue_bytes_per_parcel"] + ent["bytes_per_parcel_step"] * steps)
    if ent.get("steps_per_launch", 1) != steps:
        return None
    return ent["bytes_per_parcel_step"] * P * steps


def reduce_M_bytes_per_node(L):
    """Algorithmic HBM bytes of reduce_M per node: the n complex columns of the
    (L+2)^2/4 reduced rows read once (16 B each), four (rn, rn) FP64 outputs written."""
    n, rn = (L + 1) * (L + 2) // 2, (L + 2) ** 2 // 4
    return rn * n * 16 + 4 * rn * rn * 8


def bench_export(a):
    """Eulerian operator export: M_LROT of every mesh node built on the device and
    reduced to (Mrr, Mri, Mir, Mii) (src/specfabpy/fenics/CPO.py:200 with
    src/reducedform.f90:76-120), 2^14 nodes at L = 12."""
    import numpy as np
    import torch
    from specfab_amd import batch as B
    P, L = 1 << 14, 12
    rng = np.random.default_rng(9)
    e = rng.standard_normal((P, 3, 3))
    w = rng.standard_normal((P, 3, 3))
    eps = torch.from_numpy(e + e.transpose(0, 2, 1)).cuda()
    omg = torch.from_numpy(w - w.transpose(0, 2, 1)).cuda()
    M = B.M_LROT_batch(L, eps, omg, 1.0, 0.0)
    for _ in range(a.warmup):
        B.M_LROT_batch(L, eps, omg, 1.0, 0.0)
        B.reduce_M_batch(L, M)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(a.steps):
        B.M_LROT_batch(L, eps, omg, 1.0, 0.0)
    ev[1].record()
    for _ in range(a.steps):
        out = B.reduce_M_batch(L, M)
    ev[2].record()
    torch.cuda.synchronize()
    ms_lrot = ev[0].elapsed_time(ev[1]) / a.steps
    ms_red = ev[1].elapsed_time(ev[2]) / a.steps
    del out
    # the fused path: the reduced blocks straight from the coupling tables
    for _ in range(max(1, a.warmup)):
        B.M_LROT_reduced_batch(L, eps, omg, 1.0, 0.0)
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(a.steps):
        out = B.M_LROT_reduced_batch(L, eps, omg, 1.0, 0.0)
    ev[1].record()
    torch.cuda.synchronize()
    ms_fused = ev[0].elapsed_time(ev[1]) / a.steps
    del out
    rn = (L + 2) ** 2 // 4
    per_fused = 4 * rn * rn * 8 + 2 * 9 * 8     # the four blocks written once, eps and omg read
    ach_fused = per_fused * P / (ms_fused * 1e-3) / 1e9
    n = (L + 1) * (L + 2) // 2
    per_dense = n * n * 16 + 2 * 9 * 8          # M written once, eps and omg read
    per = reduce_M_bytes_per_node(L)
    ach = per_dense * P / (ms_lrot * 1e-3) / 1e9
    ach_red = per * P / (ms_red * 1e-3) / 1e9
    tab = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                                      "evolve_traffic.json")))["kernels"]
    traffic = {k: tab[k]["bytes_per_parcel"] * P if k in tab and tab[k]["parcels"] == P else None
               for k in ("dense_op_kernel", "reduce_M_kernel")}
    cpu = None
    if not a.no_cpu:
        from oracle import ref
        if ref.available():
            r = ref.RefLib()
            r.init(L)
            k, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < 10.0:
                r.reduce_M(r.M_LROT(e[k % P] + e[k % P].T, w[k % P] - w[k % P].T, 1.0, 0.0))
                k += 1
            cpu = {"value": k / (time.perf_counter() - t0), "unit": "nodes/s", "cores": 1, "kind": "reference",
                   "sample": "%d nodes, M_LROT + reduce_M of the reference FP64 Fortran (oracle/_ref), 10 s" % k}
    out = {"metric": "nodes/s, Eulerian operator export (M_LROT + reduce_M per node)",
           "value": P / (ms_fused * 1e-3), "unit": "nodes/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms_fused, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded)",
           "config": {"workload": "Eulerian operator export: 2^14 nodes, L=12 (91 coeffs, 49 reduced), M_LROT "
                                  "reduced to Mrr/Mri/Mir/Mii in one pass (the nodes' coupling weights, then the blocks as a "
                                  "sparse linear map of them)", "nodes": P},
           "ms_fused": ms_fused, "ms_M_LROT": ms_lrot, "ms_reduce_M": ms_red,
           "chain_nodes_per_s": P / ((ms_lrot + ms_red) * 1e-3),
           "roofline": {"bound": "hbm", "achieved": ach_fused, "peak": 8000.0, "unit": "GB/s",
                        "frac": ach_fused / 8000.0, "traffic": None,
                        "kernel": "dense_weights_kernel + reduced_map_kernel (both launches timed)",
                        "bytes_per_node": per_fused},
           "roofline_dense_op": {"bound": "hbm", "achieved": ach, "peak": 8000.0, "unit": "GB/s", "frac": ach / 8000.0,
                                 "traffic": traffic["dense_op_kernel"], "kernel": "dense_op_kernel",
                                 "bytes_per_node": per_dense},
           "roofline_reduce_M": {"bound": "hbm", "achieved": ach_red, "peak": 8000.0, "unit": "GB/s",
                                 "frac": ach_red / 8000.0, "traffic": traffic["reduce_M_kernel"],
                                 "kernel": "reduce_M_kernel", "bytes_per_node": per},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def homogenization_block(fb, procs, precision, more_steps):
    """The second half of the metric, on the field after 200 RK4 steps (the
    timed steps plus `more_steps` untimed ones): Eij (Cartesian, ice linear
    Sachs/Taylor mix) + elastic Vi along z per parcel by the plain FP64
    homogenization and by the learned one (network + closed-form Sachs term,
    envelope / error-bound fallback and a 1 % audit, stream-ordered), both
    timed with HIP events; the learned outputs of every parcel are then
    compared with the plain ones (worst parcel)."""
    import torch
    from specfab_amd import surrogate as S
    try:
        if more_steps > 0:
            fb.evolve(more_steps, 0.5, scheme="rk4", procs=procs, precision=precision)
        net = S.Surrogate().load()
        hp = S.ice_params()
        g = (hp.Eij_grain[0], hp.Eij_grain[1])

        def plain():
            E = fb.Eij(g, hp.alpha, hp.n_grain, frame="xyz")
            V = fb.Vi_elastic(hp.vi_alpha, list(hp.Lame), hp.rho, [0.0], [0.0])
            return E, V

        def learned():
            return net.homogenize(fb, hp, audit_frac=0.01, seed=3, stats=False)

        E0, E1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ms = {}
        for tag, fn in (("plain", plain), ("learned", learned)):
            fn()
            torch.cuda.synchronize()
            E0.record()
            for _ in range(5):
                fn()
            E1.record()
            torch.cuda.synchronize()
            ms[tag] = E0.elapsed_time(E1) / 5
        El, Vl, st = net.homogenize(fb, hp, audit_frac=1.0, seed=3)
        Ep, Vp = plain()
        Vs = torch.sort(Vp[:, 0, :], dim=1).values
        eE = ((El - Ep).abs() / Ep.abs()).max(1).values
        eV = ((Vl - Vs).abs() / Vs.abs()).max(1).values
        P = fb.P
        return {"field": "the timed field after 200 RK4 steps (%d parcels, L=%d)" % (P, fb.L),
                "plain_evals_per_s": P / (ms["plain"] * 1e-3), "learned_evals_per_s": P / (ms["learned"] * 1e-3),
                "learned_vs_plain": ms["plain"] / ms["learned"], "ms_plain": ms["plain"], "ms_learned": ms["learned"],
                "eval": "Eij (6, Cartesian, ice linear Sachs/Taylor mix alpha=0.0125) + elastic Vi (3, along z, "
                        "Bennett 1968 ice, Hill average) per parcel",
                "learned_dtype": "bf16 MFMA network (fp32 accumulate) for the Taylor term and Vi, fp64 Sachs term",
                "eij_rel_err_max": eE.max().item(), "eij_rel_err_mean": eE.mean().item(),
                "vi_rel_err_max": eV.max().item(), "tolerance": hp.eij_tol,
                "n_fallback": st["n_fallback"], "n_bound": st["n_bound"]}
    except Exception as e:   # the stepping metric above stands on its own
        return {"error": "%s: %s" % (type(e).__name__, e)}


# FP64 flop per parcel of eij_sachs3_kernel: the a4 and eighth-moment maps
# (2 x 2,250), 12 evaluations of the n' = 3 Sachs contraction (6 stress
# states, the parcel's and the isotropic state's), the a2 maps and the
# Jacobi eigenframe, as SQ_INSTS_VALU_FLOPS_FP64 counts them (x 64 lanes per
# parcel: 24,064); ev_D4_kernel likewise (5,632)
FLOPS_SACHS3 = 24064
FLOPS_EVD4 = 5632


def bench_nonlinear(a):
    """Nonlinear-grain homogenization on the 2^20-parcel L = 12 field (the
    headline field's synthetic states, unevolved): Eij with n' = 3 grains in
    each parcel's a2 eigenframe (Sachs, src/homogenizations.f90:93-103, and
    the Sachs/Taylor mix alpha = 0.0125), <D> with n_RSS = 4 and E_CAFFE
    (n_grain = 3) of per-parcel strain rates."""
    import torch
    from specfab_amd import batch as B
    dev = torch.device("cuda:0")
    P, L = a.parcels, 12
    from specfab_amd import _lib
    G, gam0, x = make_field(P, L, 1, dev)
    fb = B.FabricBatch.__new__(B.FabricBatch)
    fb.ctx = _lib.default_context()
    fb.L, fb.n, fb.dev, fb.P = L, (L + 1) * (L + 2) // 2, dev, P
    fb.x, fb.ugrad, fb.tau, fb.gam0, fb.lam, fb.steps = x, G, None, gam0, None, None
    fb.status = torch.zeros(P, dtype=torch.int32, device=dev)
    D = ((G + G.transpose(1, 2)) / 2).contiguous()
    Eg = (1.0, 1e3)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(fn):
        for _ in range(max(1, a.warmup)):
            fn()
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(a.steps):
            fn()
        ev[1].record()
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / a.steps

    ms = {"eij3_sachs": timed(lambda: fb.Eij(Eg, 0.0, 3, frame="eig")),
          "eij3_mix": timed(lambda: fb.Eij(Eg, 0.0125, 3, frame="eig")),
          "ev_D4": timed(lambda: fb.deformability(D, 4)),
          "caffe3": timed(lambda: fb.E_CAFFE(D, 0.1, 10.0, 3))}
    assert int(fb.status.abs().sum()) == 0
    ach = FLOPS_SACHS3 * P / (ms["eij3_sachs"] * 1e-3) / 1e12
    out = {"metric": "Eij evals/s with n'=3 grains (Sachs, a2 eigenframe), FP64",
           "value": P / (ms["eij3_sachs"] * 1e-3), "unit": "parcel-evals/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms["eij3_sachs"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded)",
           "config": {"workload": "nonlinear grains on the 2^20-parcel L=12 field: Eij n'=3 (eigenframe), "
                                  "<D> n_RSS=4, E_CAFFE n'=3", "parcels": P},
           "ms": ms, "evals_per_s": {k: P / (v * 1e-3) for k, v in ms.items()},
           "roofline": {"bound": "valu", "achieved": ach, "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
                        "frac": ach / PEAK_TFLOPS["fp64"], "traffic": None, "kernel": "eij_sachs3_kernel",
                        "flops_per_parcel": FLOPS_SACHS3},
           "roofline_ev_D4": {"bound": "valu", "achieved": FLOPS_EVD4 * P / (ms["ev_D4"] * 1e-3) / 1e12,
                              "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
                              "frac": FLOPS_EVD4 * P / (ms["ev_D4"] * 1e-3) / 1e12 / PEAK_TFLOPS["fp64"],
                              "kernel": "ev_D4_kernel", "flops_per_parcel": FLOPS_EVD4},
           "cpu_baseline": None}
    if not a.no_cpu:
        from oracle import ref
        if ref.available():
            import numpy as np
            r = ref.RefLib()
            r.init(L)
            X = (x[:, :256].T.cpu().numpy())
            from specfab_amd import specfabpy as sf
            nl = sf.init(L)[1]
            # states as complex nlm for the reference
            lm = [(l, m) for l in range(0, L + 1, 2) for m in range(-l, l + 1)]
            Z = np.zeros((X.shape[0], nl), complex)
            q = 0
            pos = {v: i for i, v in enumerate(lm)}
            for l in range(0, L + 1, 2):
                Z[:, pos[(l, 0)]] = X[:, q]
                q += 1
                for m in range(1, l + 1):
                    z = X[:, q] + 1j * X[:, q + 1]
                    Z[:, pos[(l, m)]] = z
                    Z[:, pos[(l, -m)]] = (-1) ** m * np.conj(z)
                    q += 2
            k, t0 = 0, time.perf_counter()
            I = np.eye(3)
            while time.perf_counter() - t0 < 10.0:
                r.Eij_tranisotropic(Z[k % len(Z)], I[0], I[1], I[2], Eg, 0.0, 3)
                k += 1
            el = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": k / el, "unit": "parcel-evals/s", "cores": 1, "kind": "reference",
                                   "sample": "%d Eij_tranisotropic(n'=3) calls of the reference FP64 Fortran "
                                             "(oracle/_ref), %.1f s" % (k, el)}
    print(json.dumps(out), flush=True)


# eij_ortho_kernel (v given, Cartesian frame) per parcel: algorithmic HBM
# bytes (three l <= 4 blocks read, 6 FP64 written) and FP64 flop as
# SQ_INSTS_VALU_FLOPS_FP64 counts them (x 64 lanes per parcel); vi_ortho_kernel
# (v given) FP64 flop per (parcel, direction) likewise
BYTES_ORTHO_EIJ = 3 * 15 * 8 + 6 * 8
FLOPS_ORTHO_EIJ = 1569
FLOPS_ORTHO_VI = 3879


def bench_ortho(a):
    """Orthotropic (olivine) grains on 2^20 parcels, each with the l <= 4
    distributions of its three slip-system axes (b, n, v; seeded perturbed
    isotropic states): Eij_orthotropic in the Cartesian frame
    (src/enhancementfactors.f90:134-188) with v given and v = b x n derived,
    Vi_elastic_orthotropic (src/wavepropagation.f90:41-68) along six
    directions, Abramson (1997) olivine Lame parameters."""
    import numpy as np
    import torch
    from specfab_amd import batch as B
    from specfab_amd import workloads as W
    from specfab_amd import specfabpy as sf
    P = a.parcels
    rng = np.random.default_rng(21)
    bs = [B.FabricBatch(4, nlm=W.perturbed_isotropic(rng, P, 4, sigma=0.15)) for _ in range(3)]
    Eg = (1.5, 1.0, 0.8, 1.2, 2.0, 10.0)
    Cij = np.array([320.5e9, 196.5e9, 233.5e9, 64.0e9, 77.0e9, 78.7e9, 76.8e9, 71.6e9, 68.15e9])
    Lame, rho = sf.Cij_to_Lame_orthotropic(Cij), 3355.0
    th = np.array([0.0, 0.3, 1.2, np.pi / 2, np.pi / 2, np.deg2rad(70)])
    ph = np.array([0.0, 0.0, 0.7, 0.0, 2.0, np.deg2rad(10)])
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def timed(fn):
        for _ in range(max(1, a.warmup)):
            fn()
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(a.steps):
            fn()
        ev[1].record()
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / a.steps

    ms = {"eij_given": timed(lambda: B.eij_orthotropic(bs[0], bs[1], bs[2], Eg)),
          "eij_derived": timed(lambda: B.eij_orthotropic(bs[0], bs[1], None, Eg)),
          "vi_given": timed(lambda: B.vi_elastic_orthotropic(bs[0], bs[1], bs[2], Lame, rho, th, ph)),
          "vi_derived": timed(lambda: B.vi_elastic_orthotropic(bs[0], bs[1], None, Lame, rho, th, ph))}
    E = B.eij_orthotropic(bs[0], bs[1], bs[2], Eg)
    assert bool(torch.isfinite(E).all())
    rate = {k: P * (len(th) if k.startswith("vi") else 1) / (v * 1e-3) for k, v in ms.items()}
    # the Eij kernel's algorithmic bytes need longer at 8 TB/s than its flop at
    # 78.6 TFLOP/s: HBM-bound
    ach = BYTES_ORTHO_EIJ * P / (ms["eij_given"] * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": ach, "peak": 8000.0, "unit": "GB/s", "frac": ach / 8000.0,
            "traffic": measured_traffic("eij_ortho_kernel", P, 1), "kernel": "eij_ortho_kernel",
            "bytes_per_parcel": BYTES_ORTHO_EIJ, "flops_per_parcel": FLOPS_ORTHO_EIJ,
            "fp64_frac": FLOPS_ORTHO_EIJ * P / (ms["eij_given"] * 1e-3) / 1e12 / PEAK_TFLOPS["fp64"]}
    out = {"metric": "Eij_orthotropic evals/s (olivine, Sachs, v given), FP64", "value": rate["eij_given"],
           "unit": "parcel-evals/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": ms["eij_given"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic (seeded perturbed isotropic l<=4 distributions)",
           "config": {"workload": "orthotropic grains: 2^20 parcels x 3 axis distributions (L=4), Eij "
                                  "(Cartesian frame, v given / derived), Vi along 6 directions", "parcels": P},
           "ms": ms, "evals_per_s": rate, "roofline": roof, "cpu_baseline": None}
    if FLOPS_ORTHO_VI:
        ach = FLOPS_ORTHO_VI * P * len(th) / (ms["vi_given"] * 1e-3) / 1e12
        out["roofline_vi"] = {"bound": "fp64 valu", "achieved": ach, "peak": PEAK_TFLOPS["fp64"],
                              "unit": "TFLOP/s", "frac": ach / PEAK_TFLOPS["fp64"], "kernel": "vi_ortho_kernel",
                              "flops_per_item": FLOPS_ORTHO_VI,
                              "traffic": measured_traffic("vi_ortho_kernel", P, 1)}
    if not a.no_cpu:
        from oracle import ref
        if ref.available():
            r = ref.RefLib()
            r.init(4)
            Q = [W.perturbed_isotropic(np.random.default_rng(22), 64, 4, sigma=0.15) for _ in range(3)]
            I = np.eye(3)
            k, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < 10.0:
                j = k % 64
                r.Eij_orthotropic(Q[0][j], Q[1][j], Q[2][j], I[0], I[1], I[2], Eg, 0.0, 1)
                k += 1
            el = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": k / el, "unit": "parcel-evals/s", "cores": 1, "kind": "reference",
                                   "sample": "%d Eij_orthotropic calls (v given) of the reference FP64 Fortran "
                                             "(oracle/_ref), %.1f s" % (k, el)}
    print(json.dumps(out), flush=True)


def bench_discrete(a):
    """Discrete-grain (DDM) path of the reference's SDM-vs-DDM harness
    (tests/SDM-DDM-performance/SDM-DDM-performance.py:24-73): ri_LROT
    (src/dynamics.f90:112-137) of 2^20 grain c-axes over Nt = 2000 Euler steps
    of pure shear along z (T = 1 / ln 2, t = 4, dt = t / Nt), the whole
    trajectory (Nt, N, 3) written to HBM; then ri_to_nlm (src/moments.f90:98-131)
    of the end state grouped as 1024 ensembles of 1024 grains at L = 6."""
    import numpy as np
    import torch
    from specfab_amd import batch as B
    from specfab_amd import specfabpy as sf
    N, Nt, E, G = 1 << 20, 2000, 1024, 1024
    T = 1 / math.log(2)
    ugrad = sf.pureshear_ugrad(2, 0.0, T)
    D, W = (ugrad + ugrad.T) / 2, (ugrad - ugrad.T) / 2
    dt = 4.0 / Nt
    rng = np.random.default_rng(12)
    r0 = rng.standard_normal((N, 3))
    r0 /= np.linalg.norm(r0, axis=1, keepdims=True)
    dev = torch.device("cuda:0")
    r0d = torch.from_numpy(r0).to(dev)
    Dn = torch.from_numpy(np.repeat(D[None], Nt, axis=0)).to(dev)
    Wn = torch.from_numpy(np.repeat(W[None], Nt, axis=0)).to(dev)
    wi = torch.full((G,), 1.0 / G, dtype=torch.float64, device=dev)
    fb = B.FabricBatch(6, P=E)
    ri = None
    for _ in range(a.warmup):
        ri = B.ri_LROT_batch(r0d, dt, Nt, Dn, Wn, 1.0)
        fb.set_grains(ri[-1, :E * G].reshape(E, G, 3), wi)
    del ri
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ms_lrot, ms_r2n = 0.0, 0.0
    for _ in range(a.steps):
        ev[0].record()
        ri = B.ri_LROT_batch(r0d, dt, Nt, Dn, Wn, 1.0)
        ev[1].record()
        fb.set_grains(ri[-1, :E * G].reshape(E, G, 3), wi)
        ev[2].record()
        torch.cuda.synchronize()
        ms_lrot += ev[0].elapsed_time(ev[1])
        ms_r2n += ev[1].elapsed_time(ev[2])
    # the same trajectory buffer written by a plain fill: the device's write ceiling for this size
    ri.fill_(0.0)
    ev[0].record()
    ri.fill_(1.0)
    ev[1].record()
    torch.cuda.synchronize()
    fill_gbs = ri.numel() * 8 / (ev[0].elapsed_time(ev[1]) * 1e-3) / 1e9
    del ri
    ms_lrot /= a.steps
    ms_r2n /= a.steps
    gsteps = N * (Nt - 1)
    per_gs = 24                                   # one (x, y, z) FP64 axis written per grain-step
    ach = (per_gs * gsteps + 24 * N) / (ms_lrot * 1e-3) / 1e9
    per_grain = 24                                # one axis read per grain (weights shared, cached)
    ach_r2n = (per_grain * E * G + E * 28 * 8) / (ms_r2n * 1e-3) / 1e9
    tab = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                                      "evolve_traffic.json")))["kernels"]
    tr = tab.get("ri_lrot_kernel")
    traffic = tr["bytes_per_parcel"] * gsteps if tr and tr.get("parcels") == N else None
    cpu = None
    if not a.no_cpu:
        from oracle import ref
        if ref.available():
            r = ref.RefLib()
            Ns, Ds, Ws = 1000, np.repeat(D[None], Nt, axis=0), np.repeat(W[None], Nt, axis=0)
            k, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < 10.0:
                r.ri_LROT(r0[:Ns], dt, Nt, Ds, Ws, 1.0)
                k += 1
            el = time.perf_counter() - t0
            cpu = {"value": k * Ns * (Nt - 1) / el, "unit": "grain-steps/s", "cores": 1, "kind": "reference",
                   "sample": "%d calls of ri_LROT on 1000 grains x 2000 steps, reference FP64 Fortran "
                             "(oracle/_ref), %.1f s" % (k, el)}
    out = {"metric": "grain-steps/s, discrete lattice rotation (ri_LROT, whole trajectory written)",
           "value": gsteps / (ms_lrot * 1e-3), "unit": "grain-steps/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms_lrot, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded random unit axes)",
           "config": {"workload": "DDM: ri_LROT of 2^20 grains, Nt=2000 Euler steps of pure shear along z "
                                  "(T=1/ln2, t=4), trajectory (2000, 2^20, 3) in HBM; ri_to_nlm L=6 of "
                                  "the end state as 1024 ensembles x 1024 grains", "grains": N, "Nt": Nt},
           "ms_ri_to_nlm": ms_r2n, "ensembles_per_s": E / (ms_r2n * 1e-3),
           "roofline": {"bound": "hbm", "achieved": ach, "peak": 8000.0, "unit": "GB/s", "frac": ach / 8000.0,
                        "traffic": traffic, "kernel": "ri_lrot_kernel", "bytes_per_grain_step": per_gs,
                        "fill_GBps": fill_gbs, "frac_of_fill": ach / fill_gbs},
           "roofline_ri_to_nlm": {"bound": "hbm", "achieved": ach_r2n, "peak": 8000.0, "unit": "GB/s",
                                  "frac": ach_r2n / 8000.0, "kernel": "ri_moments_kernel",
                                  "bytes_per_grain": per_grain},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--parcels", type=int, default=1 << 20)
    ap.add_argument("--L", type=int, default=12)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32c"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="field",
                    choices=["field", "uneven", "surrogate", "export", "discrete", "nonlinear", "ortho"],
                    help="field: the L=12 Eulerian field (default, the headline); uneven: 2^18 parcels with mixed "
                         "L and step counts in one graph launch; surrogate: in-loop homogenization, plain vs "
                         "learned; export: operator export; discrete: ri_LROT / ri_to_nlm of 2^20 grains; "
                         "nonlinear: n'=3 Eij, <D> and E_CAFFE on the field; ortho: olivine Eij / Vi")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (scale reference for FETCH_SIZE / WRITE_SIZE)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU; parcels shard across GPUs without communication, but the "
              "multi-GPU run is not part of this benchmark (--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    if a.copy_probe:
        n = 1 << 27
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print("copy probe: 3 copies of %d bytes" % (n * 8), flush=True)
        return
    from specfab_amd import batch as B
    from specfab_amd import _lib

    if a.workload == "uneven":
        return bench_uneven(a)
    if a.workload == "surrogate":
        return bench_surrogate(a)
    if a.workload == "export":
        return bench_export(a)
    if a.workload == "discrete":
        return bench_discrete(a)
    if a.workload == "nonlinear":
        return bench_nonlinear(a)
    if a.workload == "ortho":
        return bench_ortho(a)
    dev = torch.device("cuda:0")
    P, L = a.parcels, a.L
    G, gam0, x = make_field(P, L, 1, dev)
    fb = B.FabricBatch.__new__(B.FabricBatch)
    fb.ctx = _lib.default_context()
    fb.L, fb.n, fb.dev, fb.P = L, (L + 1) * (L + 2) // 2, dev, P
    fb.x, fb.ugrad, fb.tau, fb.gam0, fb.lam, fb.steps = x, G, None, gam0, None, None
    fb.status = torch.zeros(P, dtype=torch.int32, device=dev)
    procs = ("LROT", "REG", "DDRX")
    # autotune on the full field, so that every launch of the chosen kernel
    # (also in a rocprofv3 trace of this command) has the timed size
    variant, variant_ms = fb.autotune(a.precision, P)
    # the time loop as a caller runs it: all steps in one call (one launch of
    # the stepper, whose lanes loop over the steps); warmup likewise
    if a.warmup:
        fb.evolve(a.warmup, 0.5, scheme="rk4", procs=procs, precision=a.precision)
    torch.cuda.synchronize()
    # the stepping kernel is bracketed by HIP events inside the library, on the
    # stream it is launched on (torch's current stream)
    lib = _lib.lib()
    _lib.check(lib.sfg_ctx_set_timing(fb.ctx.handle, 1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fb.evolve(a.steps, 0.5, scheme="rk4", procs=procs, precision=a.precision)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    import ctypes
    tot, cnt = ctypes.c_double(0), ctypes.c_int(0)
    _lib.check(lib.sfg_ctx_read_timing(fb.ctx.handle, ctypes.byref(tot), ctypes.byref(cnt)))
    _lib.check(lib.sfg_ctx_set_timing(fb.ctx.handle, 0))
    launch_ms = tot.value / max(1, cnt.value)   # one launch of a.steps steps
    value = P * a.steps / wall
    nonfinite = int((fb.status != 0).sum().item())
    flops = flops_per_parcel_step(L) * P * a.steps
    achieved = flops / (launch_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[a.precision]
    kname = "evolve_fast_L%d_%s_v%d" % (L, "f64" if a.precision == "fp64" else "f32", variant)
    traffic = measured_traffic(kname, P, a.steps)
    # Eij of the evolved field (eigenframe, ice linear Sachs/Taylor mix)
    E0, E1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fb.Eij((1.0, 1e3), 0.0125, 1)
    E0.record()
    for _ in range(3):
        fb.Eij((1.0, 1e3), 0.0125, 1)
    E1.record()
    torch.cuda.synchronize()
    eij_rate = 3 * P / (E0.elapsed_time(E1) * 1e-3)
    # elastic velocities of the evolved field: Hill average (alpha = 0.5),
    # Bennett (1968) ice, rho = 917, k along z
    lame = [7.15e9, 3.455e9, 5.88e9 / 7.15e9, 3.06e9 / 3.455e9, 15.24e9 / 14.06e9]
    fb.Vi_elastic(0.5, lame, 917.0, [0.0], [0.0])
    E0.record()
    for _ in range(3):
        fb.Vi_elastic(0.5, lame, 917.0, [0.0], [0.0])
    E1.record()
    torch.cuda.synchronize()
    vi_rate = 3 * P / (E0.elapsed_time(E1) * 1e-3)
    # the rest of the 200-step run as one launch, timed like the stepping
    # launch above: the sustained rate (the short timed launch starts right
    # after the autotuner's host-side teardown, while the GPU clock is still
    # rising from that idle gap)
    more = 200 - a.warmup - a.steps
    sustained = None
    if more > 0:
        _lib.check(lib.sfg_ctx_set_timing(fb.ctx.handle, 1))
        fb.evolve(more, 0.5, scheme="rk4", procs=procs, precision=a.precision)
        torch.cuda.synchronize()
        _lib.check(lib.sfg_ctx_read_timing(fb.ctx.handle, ctypes.byref(tot), ctypes.byref(cnt)))
        _lib.check(lib.sfg_ctx_set_timing(fb.ctx.handle, 0))
        sms = tot.value / max(1, cnt.value)
        sach = flops_per_parcel_step(L) * P * more / (sms * 1e-3) / 1e12
        sustained = {"steps_per_launch": more, "launch_ms": sms, "ms_per_step": sms / more,
                     "parcel_steps_per_s": P * more / (sms * 1e-3), "achieved": sach, "frac": sach / peak,
                     "traffic": measured_traffic(kname, P, more)}
    homog = homogenization_block(fb, procs, a.precision, 0)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "parcel-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32+compensated",
        "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": "Eulerian CPO field: %d parcels, L=%d (%d coeffs), LROT+REG+DDRX, RK4 dt=0.5"
                               % (P, L, fb.n), "parcels": P, "L": L, "scheme": "rk4", "dt": 0.5,
                   "procs": list(procs), "full_run_steps": 200, "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic,
                     "compute": "%s vector ALU (sparse FMA; MI355X %s peak %.1f TFLOP/s)"
                                % ("fp64" if a.precision == "fp64" else "fp32", a.precision, peak),
                     "kernel": "%s (one launch: %d RK4 steps of the field)" % (kname, a.steps),
                     "flop_per_parcel_step": flops_per_parcel_step(L), "steps_per_launch": a.steps,
                     "launch_ms": launch_ms, "sustained": sustained},
        "autotune": {"variant": variant, "ms_per_step_by_variant": variant_ms, "probe_parcels": P},
        "eij_evals_per_s": eij_rate,
        "vi_elastic_evals_per_s": vi_rate,
        "nonfinite_parcels": nonfinite,
        "homogenization": homog,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
