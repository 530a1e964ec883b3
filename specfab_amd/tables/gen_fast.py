#!/usr/bin/env python3
"""Generate the register-resident stepping kernels (HIP, gfx950).

For each truncation L in FAST_L and each precision, this writes
csrc/generated/sfg_fast_L{L}_{f64|f32}.hip containing one kernel that steps
64 parcels per wavefront (one parcel per lane) through all its time steps:

  * the stage input x (n real coefficients) lives in VGPRs as n scalars;
  * every coupling value is compiled into the instruction stream as a literal
    constant: FP32 terms become single v_fmamk_f32 / v_fmac_f32 instructions
    with a 32-bit literal, FP64 terms a v_fma_f64 whose 64-bit constant is
    materialised by two s_mov_b32 on the scalar unit (which co-issue with
    another wave's vector FMAs), so the operator costs no memory traffic and
    no scalar-cache latency;
  * the 23 per-parcel operator parameters theta_p sit in LDS and are read
    once per (row chunk, parameter);
  * per output row a the kernel evaluates
        f_a = sum_p theta_p (sum_b R_p[a][b] x_b) + (s_l(a) - Gamma0 <D>(x)) x_a
    and applies the Euler / RK4 update with the state x0 (FP64, HBM), the RK
    increment accumulator YS and the next-stage input XN (SoA buffers);
    Euler writes the new state directly, the first RK4 stage does not read
    YS and the last one writes only the new state;
  * __launch_bounds__ asks for WAVES waves per SIMD so that a second wave's
    vector work overlaps the scalar constant moves and memory waits.

The sparsity structure and the values come from a dump of the host builder
(specfab_amd/tables/dump_basis.cpp), i.e. the same term array that the
generic stepper reads at run time (sfg::real_operator(L).terms).

FP32 variant ("FP32 stepping, compensated state update"): the operator
application and the RK increments run in FP32 (constants rounded to float);
the state x0 is kept and updated in FP64 (x1 = x0 + sum of FP32 increments).

Usage: gen_fast.py OUTDIR
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
FAST_L = {"f64": (4, 6, 8, 10, 12, 14, 16, 18, 20), "f32": (4, 6, 8, 10, 12, 14, 16, 18, 20)}

# Kernel variants (the library's autotuner times them and keeps the fastest):
#   mode    "lit": coupling values as literal constants in the code, the RK
#           increment sum and next-stage input in SoA buffers ("reg", "win",
#           "sreg" below); coupling values read through the constant address
#           space (scalar loads) measured 3x slower than literals and were
#           dropped
#   chunk   output rows accumulated together
#   waves   waves per SIMD requested from the register allocator
#   bfree   branch-free RK update (stage flags folded into buffer offsets)
#   fence   scheduling fence after every (row, parameter) group
#   block   threads per workgroup (64 = one wave; 256: four waves dispatched
#           together)
#   sync    workgroup barrier at every row chunk (uniform step counts only):
#           the waves of a workgroup then walk the ~100 KB stage code in
#           lockstep and share its instruction-cache lines
#   xn      ("reg" mode) where the next-stage input lives: "reg" (registers,
#           with an FP32 copy of the step-start state), "mem" (XN buffer) or
#           "roll" (registers: a row's next-stage value replaces its stage
#           input as soon as no later row chunk of the stage reads it, so only
#           about two l-levels of rows are held twice)
#   ahead   ("reg" mode) the step-start state rows of row chunk c + 1 loaded
#           at the top of chunk c, before chunk c's stores: waiting for them
#           then does not wait for the stores just issued (one memory counter
#           covers loads and stores)
V = lambda mode, chunk, waves, bfree, fence, block=64, sync=False: dict(
    mode=mode, chunk=chunk, waves=waves, bfree=bfree, fence=fence, block=block, sync=sync)
#   split   ("reg" mode) the last stage as a separate copy of the stage code
#           (else one stage loop that tests the stage kind per row chunk)
R = lambda chunk, waves, xn, split=True, block=64, ahead=False: dict(
    mode="reg", chunk=chunk, waves=waves, xn=xn, split=split, block=block, ahead=ahead)
#   window  ("win" mode) stage-input columns held in registers per row chunk
#   split   ("win" mode) waves that share one group of 64 parcels, each
#           evaluating a contiguous share of the row chunks (balanced by term
#           count), with a workgroup barrier after every stage: a bucket of
#           few parcels then still runs several waves per SIMD
Wn = lambda window, chunk, waves, split=1: dict(mode="win", window=window, chunk=chunk, waves=waves,
                                                block=64 * split, split=split)
#   sreg    split register stepper (FP32): `split` waves per 64-parcel group,
#           each holding the whole stage input and the increment sums of its
#           share of the rows; stage inputs exchanged through LDS
Sr = lambda chunk, waves, split=2: dict(mode="sreg", chunk=chunk, waves=waves, block=64 * split, split=split)
#   dreg    split resident stepper (FP64): `split` waves per 64-parcel group,
#           each holding the whole stage input and the increment sums and
#           step-start state of its share of the rows (no HBM traffic per
#           step); stage inputs exchanged through LDS
#   gb      ("dreg" mode) (row chunk, parameter) groups per basic block
#   cols    ("dreg" mode) a wave holds only the stage-input columns its rows
#           read, so that eight waves (two per SIMD, 256 registers each) share
#           a parcel group and one wave's scalar constant moves issue beside
#           the other's vector FMAs
#   lwin    LDS window stepper (FP64, L >= 14): `split` waves per 64-parcel
#           group, the group's stage input in LDS, each wave's rows evaluated
#           in column windows loaded from LDS
Lw = lambda window, chunk, split=4, gb=2: dict(mode="lwin", window=window, chunk=chunk, waves=1, block=64 * split,
                                              split=split, gb=gb)
Dr = lambda chunk, split=2, gb=2, cols=False: dict(mode="dreg", chunk=chunk, waves=2 if cols else 1,
                                                  block=64 * split, split=split, gb=gb, cols=cols)
VARIANTS = {
    ("f64", "tuned"): [V("lit", 8, 2, False, False), V("lit", 8, 1, False, False),
                       V("lit", 8, 2, False, False, 256, True), R(8, 1, "mem", False), R(4, 1, "mem", False),
                       R(4, 1, "roll", False, ahead=True), Dr(4), Dr(4, 4), Dr(8, 4), Dr(8, 2),
                       Dr(8, 4, gb=1), Dr(16, 4, gb=4), Dr(4, 8, cols=True), Dr(8, 8, cols=True),
                       Dr(16, 8, gb=4, cols=True), Dr(16, 8, gb=8, cols=True)],
    ("f64", "default"): [V("lit", 8, 2, False, False)],
    ("f32", "tuned"): [V("lit", 8, 2, True, False), V("lit", 8, 3, True, False), R(8, 1, "reg"),
                       R(8, 2, "mem"), R(4, 1, "reg"), R(16, 1, "reg"), Sr(8, 2)],
    ("f32", "default"): [V("lit", 8, 2, True, False)],
    ("f64", "wide"): [Wn(96, 8, 1), Wn(64, 8, 2), Wn(64, 8, 2, split=2), Lw(96, 8), Lw(64, 8), Lw(128, 8),
                      Lw(64, 4, split=8), Lw(48, 4, split=8)],
    ("f32", "wide"): [Wn(128, 8, 1), Wn(96, 8, 2), Wn(96, 8, 2, split=2), Lw(128, 8), Lw(96, 8, split=8),
                      Lw(64, 8, split=8)],
}
WIDE_L = {"f64": (14, 16, 18, 20), "f32": (18, 20)}
TUNED_L = (8, 12)


def dump(L, exe):
    out = subprocess.check_output([exe, str(L)]).decode().split("\n")
    L_, n, nrows, ngroups, nterms = map(int, out[0].split())
    rows = []
    i = 1
    while i < len(out) and out[i]:
        tok = out[i].split()
        assert tok[0] == "R"
        l, ng = int(tok[1]), int(tok[2])
        i += 1
        groups = []
        for _ in range(ng):
            g = out[i].split()
            assert g[0] == "G"
            p, nt = int(g[1]), int(g[2])
            i += 1
            terms = []
            for _ in range(nt):
                t = out[i].split()
                terms.append((int(t[1]), int(t[2]), float.fromhex(t[3])))
                i += 1
            groups.append((p, terms))
        rows.append((l, groups))
    assert len(rows) == n
    return n, rows


def lit(v, prec):
    """Exact C literal of a term value (FP32: rounded to float)."""
    if prec == "f64":
        return "(%s)" % float(v).hex()
    import struct
    f = struct.unpack("f", struct.pack("f", v))[0]
    return "(%sf)" % float(f).hex()


def gen_kernel(L, prec, n, rows, cfg, vid):
    T = "double" if prec == "f64" else "float"
    LD = "bld64" if prec == "f64" else "bld32"
    ST = "bst64" if prec == "f64" else "bst32"
    W = 8 if prec == "f64" else 4
    cv = (lambda e: e) if prec == "f64" else (lambda e: "(float)(%s)" % e)
    name = "evolve_fast_L%d_%s_v%d" % (L, prec, vid)
    CHUNK, BRANCHFREE, FENCE = cfg["chunk"], cfg["bfree"], cfg["fence"]
    K = lambda v, ti: lit(v, prec)
    used_p = sorted({p for (_, gs) in rows for (p, _) in gs})
    nq = min(n, 15)
    nblk = L // 2 + 1
    o = []
    w = o.append
    w("// Generated by specfab_amd/tables/gen_fast.py -- do not edit.")
    w("// Register-resident stepper, L = %d (n = %d), %s; variant %d: %s." %
      (L, n, "FP64" if prec == "f64" else "FP32 stepping, FP64 state", vid,
       ", ".join("%s=%s" % kv for kv in sorted(cfg.items()))))
    w('#include "../sfg_kernels.h"')
    w("namespace sfg {")
    w("__global__ void __launch_bounds__(%d, %d) %s(EvolveArgs a) {" % (cfg["block"], cfg["waves"], name))
    w("    const int64_t P = a.P;")
    B = cfg["block"]
    w("    const int64_t p = (int64_t)blockIdx.x * %d + threadIdx.x;" % B)
    SYNC = cfg["sync"]
    if SYNC:
        w("    // lockstep mode: with uniform step counts every lane of the workgroup")
        w("    // (also those past P) runs the whole loop nest and meets every barrier;")
        w("    // lanes past P address the buffers out of range (loads 0, stores dropped)")
        w("    const bool bsync = (a.steps == nullptr);")
        w("    const bool live = p < P;")
        w("    if (!bsync && !live) return;")
        w("    const int64_t pc = live ? p : P - 1;")
    else:
        w("    if (p >= P) return;")
        w("    const int64_t pc = p;")
    w("    // SoA rows through buffer resources: row r of an array with element")
    w("    // size w is at byte (r * P * w) [scalar] + p * w [one lane offset]")
    w("    const uint32_t P8 = (uint32_t)P * 8u, PT = (uint32_t)P * %du;" % W)
    if SYNC:
        w("    const uint32_t v8 = live ? (uint32_t)p * 8u : 0x80000000u, vT = live ? (uint32_t)p * %du : 0x80000000u;" % W)
    else:
        w("    const uint32_t v8 = (uint32_t)p * 8u, vT = (uint32_t)p * %du;" % W)
    w("    const sfg_rsrc RX = make_rsrc(a.X, P8 * %du);" % n)
    w("    const sfg_rsrc RYS = make_rsrc(a.YS, PT * %du);" % n)
    w("    const sfg_rsrc RXN = make_rsrc(a.XA, PT * %du);" % n)
    w("    const sfg_rsrc ROM = make_rsrc(a.omega, P8 * 15u);")
    w("    const sfg_rsrc RSL = make_rsrc(a.sl, P8 * %du);" % nblk)
    w("    const int nstage = a.scheme ? 4 : 1;")
    w("    const int rem = a.steps ? a.steps[pc] : a.nsteps;")
    w("    const int ns = (a.steps && a.steps_dec) ? (rem < a.nsteps ? rem : a.nsteps) : rem;")
    w("    const double h = a.dt;")
    w("    // per-parcel operator parameters in LDS ([param][lane]), read once per")
    w("    // (row chunk, parameter)")
    w("    __shared__ %s thl[%d][%d];" % (T, len(used_p), B))
    pidx = {q: i for i, q in enumerate(used_p)}
    for q in used_p:
        w("    thl[%d][threadIdx.x] = %s;" % (pidx[q], cv("a.theta[(int64_t)%d * P + pc]" % q)))
    w("    %s %s;" % (T, ", ".join("x%d" % b for b in range(n))))
    w("    for (int step = 0; step < ns; step++) {")
    for b in range(n):
        w("        x%d = %s;" % (b, cv("bld64(RX, v8, %du * P8)" % b)))
    w("        for (int s = 0; s < nstage; s++) {")
    w("            const bool first = (s == 0), last = (s == nstage - 1), euler = (nstage == 1);")
    w("            // a.zero is 0 at run time but unknown to the compiler: it keeps the")
    w("            // LDS parameter reads of different row chunks apart")
    w("            const int z = (int)a.zero * (step * 4 + s);")
    w("            const %s cy = %s;" % (T, cv("euler ? h : ((s == 0 || s == 3) ? h / 6 : h / 3)")))
    w("            const %s cx = %s;" % (T, cv("(s == 2) ? h : h / 2")))
    if BRANCHFREE:
        w("            // stage flags as lane offsets: 0x80000000 is past every buffer (whose size")
        w("            // is < 2^31), so the first stage reads YS as 0 and the last one drops its")
        w("            // YS / XN stores; the row offset stays in the scalar offset (< 2^31)")
        w("            const uint32_t vY = (first || euler) ? 0x80000000u : vT;")
        w("            const uint32_t vW = (last || euler) ? 0x80000000u : vT;")
    w("            %s gD = (%s)0;" % (T, T))
    w("            if (a.use_ddrx) {")
    terms = " + ".join("%s * x%d" % (cv("bld64(ROM, v8, %du * P8)" % q), q) for q in range(nq))
    w("                gD = (%s) / x0;" % terms)
    w("            }")
    for c0 in range(0, n, CHUNK):
        chunk = list(range(c0, min(n, c0 + CHUNK)))
        w("            {   // rows %d..%d" % (chunk[0], chunk[-1]))
        if SYNC:
            w("                if (bsync) __syncthreads();")
        w("                %s %s;" % (T, ", ".join("acc%d = (%s)0" % (r, T) for r in chunk)))
        byp = {}
        for r in chunk:
            for (pp, tl) in rows[r][1]:
                byp.setdefault(pp, []).append((r, tl))
        for pp in sorted(byp):
            w("                {")
            w("                    const %s th = thl[%d][threadIdx.x + z * %d];" % (T, pidx[pp], c0 // CHUNK + 1))
            for (r, tl) in byp[pp]:
                col, ti, v = tl[0]
                w("                    %s u%d = %s * x%d;" % (T, r, K(v, ti), col))
                for (col, ti, v) in tl[1:]:
                    w("                    u%d = fma(%s, x%d, u%d);" % (r, K(v, ti), col, r))
                w("                    acc%d = fma(th, u%d, acc%d);" % (r, r, r))
                if FENCE:
                    w("                    __builtin_amdgcn_sched_barrier(0);")
            w("                }")
        for r in chunk:
            l = rows[r][0]
            w("                {")
            w("                    const %s sdl = %s - gD;" % (T, cv("bld64(RSL, v8, %du * P8)" % (l // 2))))
            w("                    const %s f = fma(sdl, x%d, acc%d);" % (T, r, r))
            w("                    const double x0a = bld64(RX, v8, %du * P8);" % r)
            if BRANCHFREE:
                # uniform stage flags folded into the offsets: the first stage
                # reads YS beyond the buffer (returns 0), the last stage writes
                # YS / XN beyond it (dropped), other stages write X back unchanged
                w("                    const %s ys = fma(cy, f, %s(RYS, vY, %du * PT));" % (T, LD, r))
                w("                    bst64(RX, v8, %du * P8, (last || euler) ? x0a + (double)ys : x0a);" % r)
                w("                    %s(RYS, vW, %du * PT, ys);" % (ST, r))
                w("                    %s(RXN, vW, %du * PT, fma(cx, f, (%s)x0a));" % (ST, r, T))
            else:
                w("                    if (euler) {")
                w("                        bst64(RX, v8, %du * P8, x0a + (double)(cy * f));" % r)
                w("                    } else {")
                w("                        const %s ys = first ? cy * f : fma(cy, f, %s(RYS, vT, %du * PT));" % (T, LD, r))
                w("                        if (last) {")
                w("                            bst64(RX, v8, %du * P8, x0a + (double)ys);" % r)
                w("                        } else {")
                w("                            %s(RYS, vT, %du * PT, ys);" % (ST, r))
                w("                            %s(RXN, vT, %du * PT, fma(cx, f, (%s)x0a));" % (ST, r, T))
                w("                        }")
                w("                    }")
            w("                }")
        w("            }")
    w("            if (!last) {")
    for b in range(n):
        w("                x%d = %s(RXN, vT, %du * PT);" % (b, LD, b))
    w("            }")
    w("        }")
    w("    }")
    if SYNC:
        w("    if (!live) return;")
    w("    if (a.steps && a.steps_dec) a.steps[p] = rem - ns;")
    w("    if (a.status) {")
    w("        int st = 0;")
    w("        for (int r = 0; r < %d; r++) st |= isfinite(bld64(RX, v8, (uint32_t)r * P8)) ? 0 : 1;" % n)
    w("        a.status[p] |= st;")
    w("    }")
    w("}")
    w("#ifndef SFG_HOST_EMUL")
    w("hipError_t launch_%s(const EvolveArgs& a, hipStream_t s) {" % name)
    w("    hipLaunchKernelGGL(%s, dim3((unsigned)((a.P + %d) / %d)), dim3(%d), 0, s, a);" % (name, B - 1, B, B))
    w("    return hipGetLastError();")
    w("}")
    w("#endif")
    w("}  // namespace sfg")
    return "\n".join(o) + "\n"


def gen_kernel_reg(L, prec, n, rows, cfg, vid):
    """Register-state stepper ("reg" mode).

    Differences from gen_kernel: the RK4 increment sum y stays in registers
    for the whole step; with xn="reg" the next-stage input and the FP32 copy
    of the step-start state x0f are registers too, so a step touches HBM only
    to read and write the FP64 state once (last stage); with xn="mem" the
    next-stage input goes through XN and x0 is read per row chunk.  theta,
    the 15 DDRX moment weights omega and the per-l diagonal s_l are staged
    once in LDS.  Loads of a row chunk's x0 are issued at the top of the
    chunk so that their latency overlaps the chunk's FMA block.  The last RK
    stage (or the single Euler stage) is a separate copy of the stage code,
    so no per-row branch on the stage kind remains.
    """
    T = "double" if prec == "f64" else "float"
    LD = "bld64" if prec == "f64" else "bld32"
    ST = "bst64" if prec == "f64" else "bst32"
    W = 8 if prec == "f64" else 4
    cv = (lambda e: e) if prec == "f64" else (lambda e: "(float)(%s)" % e)
    name = "evolve_fast_L%d_%s_v%d" % (L, prec, vid)
    CHUNK, B = cfg["chunk"], cfg["block"]
    XREG = cfg["xn"] == "reg"
    ROLL = cfg["xn"] == "roll"
    AHEAD = cfg.get("ahead", False)
    used_p = sorted({p for (_, gs) in rows for (p, _) in gs})
    NPU = len(used_p)
    nq = min(n, 15)
    nblk = L // 2 + 1
    OMC, SLC = NPU, NPU + nq          # LDS columns of omega and s_l
    pidx = {q: i for i, q in enumerate(used_p)}
    o = []
    w = o.append
    w("// Generated by specfab_amd/tables/gen_fast.py -- do not edit.")
    w("// Register-state stepper, L = %d (n = %d), %s; variant %d: %s." %
      (L, n, "FP64" if prec == "f64" else "FP32 stepping, FP64 state", vid,
       ", ".join("%s=%s" % kv for kv in sorted(cfg.items()))))
    w('#include "../sfg_kernels.h"')
    w("namespace sfg {")
    w("__global__ void __launch_bounds__(%d, %d) %s(EvolveArgs a) {" % (B, cfg["waves"], name))
    w("    const int64_t P = a.P;")
    w("    const int64_t p = (int64_t)blockIdx.x * %d + threadIdx.x;" % B)
    w("    if (p >= P) return;")
    w("    // SoA rows through buffer resources: row r of an array with element")
    w("    // size w is at byte (r * P * w) [scalar] + p * w [one lane offset]")
    w("    const uint32_t P8 = (uint32_t)P * 8u, PT = (uint32_t)P * %du;" % W)
    w("    const uint32_t v8 = (uint32_t)p * 8u, vT = (uint32_t)p * %du;" % W)
    w("    const sfg_rsrc RX = make_rsrc(a.X, P8 * %du);" % n)
    if not XREG and not ROLL:
        w("    const sfg_rsrc RXN = make_rsrc(a.XA, PT * %du);" % n)
    w("    const int nstage = a.scheme ? 4 : 1;")
    w("    const int rem = a.steps ? a.steps[p] : a.nsteps;")
    w("    const int ns = (a.steps && a.steps_dec) ? (rem < a.nsteps ? rem : a.nsteps) : rem;")
    w("    const double h = a.dt;")
    w("    // per-parcel constants of the launch in LDS ([column][lane]): theta_p,")
    w("    // then the DDRX moment weights omega_q, then the diagonal s_l")
    w("    __shared__ %s thl[%d][%d];" % (T, NPU + nq + nblk, B))
    for q in used_p:
        w("    thl[%d][threadIdx.x] = %s;" % (pidx[q], cv("a.theta[(int64_t)%d * P + p]" % q)))
    for q in range(nq):
        w("    thl[%d][threadIdx.x] = %s;" % (OMC + q, cv("a.omega[(int64_t)%d * P + p]" % q)))
    for b in range(nblk):
        w("    thl[%d][threadIdx.x] = %s;" % (SLC + b, cv("a.sl[(int64_t)%d * P + p]" % b)))
    w("    %s %s;" % (T, ", ".join("x%d" % b for b in range(n))))
    w("    %s %s;" % (T, ", ".join("y%d = (%s)0" % (b, T) for b in range(n))))
    if XREG:
        w("    %s %s;" % (T, ", ".join("xn%d" % b for b in range(n))))
        w("    %s %s;" % (T, ", ".join("xf%d" % b for b in range(n))))
        for b in range(n):
            w("    x%d = %s; xf%d = x%d;" % (b, cv("bld64(RX, v8, %du * P8)" % b), b, b))
    chunks = [list(range(c0, min(n, c0 + CHUNK))) for c0 in range(0, n, CHUNK)]
    if AHEAD:
        w("    double %s;" % ", ".join("xo%d" % b for b in range(n)))
    commit = {}
    if ROLL:
        w("    %s %s;" % (T, ", ".join("xn%d" % b for b in range(n))))
        for b in range(n):
            w("    x%d = %s;" % (b, cv("bld64(RX, v8, %du * P8)" % b)))
        # chunk after which row b's stage input is dead: the last chunk whose
        # terms or diagonal read column b (the DDRX moment sum reads rows
        # 0..14 before the first chunk); its next-stage value is written
        # into x_b after max(that chunk, b's own chunk)
        lastread = {b: -1 for b in range(n)}
        for ci, ch in enumerate(chunks):
            for r in ch:
                lastread[r] = max(lastread[r], ci)
                for (pp, tl) in rows[r][1]:
                    for (col, ti, v) in tl:
                        lastread[col] = max(lastread[col], ci)
        for ci, ch in enumerate(chunks):
            for r in ch:
                commit.setdefault(max(ci, lastread[r]), []).append(r)
    zc = [0]

    def zk():
        zc[0] += 1
        return zc[0]

    def stage(last):
        # one stage's code; `last`: the final RK4 stage or the Euler stage;
        # None: any stage, the kind tested at run time per row chunk
        if last is None:
            w("            const bool first = (s == 0), last = (s == nstage - 1);")
            w("            const %s cy = %s;" % (T, cv("(nstage == 1) ? h : ((s == 0 || s == 3) ? h / 6 : h / 3)")))
            w("            const %s cx = %s;" % (T, cv("(s == 2) ? h : h / 2")))
        elif last:
            w("            const bool first = (nstage == 1);")
            w("            const %s cy = %s;" % (T, cv("first ? h : h / 6")))
        else:
            w("            const bool first = (s == 0);")
            w("            const %s cy = %s;" % (T, cv("(s == 0) ? h / 6 : h / 3")))
            w("            const %s cx = %s;" % (T, cv("(s == 2) ? h : h / 2")))
        w("            // a.zero is 0 at run time but unknown to the compiler: it keeps the")
        w("            // LDS reads of different row chunks and stages apart")
        w("            const int z = (int)a.zero * (step * 4 + %s);" % ("3" if last else "s"))
        dyn = last is None
        w("            %s gD = (%s)0;" % (T, T))
        w("            if (a.use_ddrx) {")
        k = zk()
        terms = " + ".join("thl[%d][threadIdx.x + z * %d] * x%d" % (OMC + q, k, q) for q in range(nq))
        w("                gD = (%s) / x0;" % terms)
        w("            }")
        loadxo = last or dyn or not XREG
        if loadxo and AHEAD:
            w("            %s" % " ".join("xo%d = bld64(RX, v8, %du * P8);" % (r, r) for r in chunks[0]))
        for ci, chunk in enumerate(chunks):
            k = zk()
            w("            {   // rows %d..%d" % (chunk[0], chunk[-1]))
            if loadxo and AHEAD:
                if ci + 1 < len(chunks):
                    w("                %s" % " ".join("xo%d = bld64(RX, v8, %du * P8);" % (r, r) for r in chunks[ci + 1]))
            elif loadxo:
                w("                const double %s;" % ", ".join("xo%d = bld64(RX, v8, %du * P8)" % (r, r) for r in chunk))
            w("                %s %s;" % (T, ", ".join("acc%d = (%s)0" % (r, T) for r in chunk)))
            byp = {}
            for r in chunk:
                for (pp, tl) in rows[r][1]:
                    byp.setdefault(pp, []).append((r, tl))
            for pp in sorted(byp):
                w("                {")
                w("                    const %s th = thl[%d][threadIdx.x + z * %d];" % (T, pidx[pp], k))
                for (r, tl) in byp[pp]:
                    col, ti, v = tl[0]
                    w("                    %s u%d = %s * x%d;" % (T, r, lit(v, prec), col))
                    for (col, ti, v) in tl[1:]:
                        w("                    u%d = fma(%s, x%d, u%d);" % (r, lit(v, prec), col, r))
                    w("                    acc%d = fma(th, u%d, acc%d);" % (r, r, r))
                w("                }")
            def epilogue(last, r):
                l = rows[r][0]
                w("                {")
                w("                    const %s f = fma(thl[%d][threadIdx.x + z * %d] - gD, x%d, acc%d);"
                  % (T, SLC + l // 2, k, r, r))
                w("                    const %s yv = fma(cy, f, first ? (%s)0 : y%d);" % (T, T, r))
                if last:
                    w("                    const double xs = xo%d + (double)yv;" % r)
                    w("                    bst64(RX, v8, %du * P8, xs);" % r)
                    if XREG:
                        w("                    xf%d = %s;" % (r, cv("xs")))
                    elif ROLL:
                        w("                    xn%d = %s;" % (r, cv("xs")))
                else:
                    w("                    y%d = yv;" % r)
                    if XREG:
                        w("                    xn%d = fma(cx, f, xf%d);" % (r, r))
                    elif ROLL:
                        w("                    xn%d = fma(cx, f, (%s)xo%d);" % (r, T, r))
                    else:
                        w("                    %s(RXN, vT, %du * PT, fma(cx, f, (%s)xo%d));" % (ST, r, T, r))
                w("                }")
            if dyn:
                w("                if (last) {")
                for r in chunk:
                    epilogue(True, r)
                w("                } else {")
                for r in chunk:
                    epilogue(False, r)
                w("                }")
            else:
                for r in chunk:
                    epilogue(last, r)
            w("            }")
            if ROLL and ci in commit:
                w("            %s" % " ".join("x%d = xn%d;" % (b, b) for b in sorted(commit[ci])))
        # next stage input
        if ROLL:
            pass
        elif dyn:
            assert not XREG
            w("            if (!last) {")
            for b in range(n):
                w("                x%d = %s(RXN, vT, %du * PT);" % (b, LD, b))
            w("            }")
        elif XREG:
            src = "xf" if last else "xn"
            w("            %s" % " ".join("x%d = %s%d;" % (b, src, b) for b in range(n)))
        elif not last:
            for b in range(n):
                w("            x%d = %s(RXN, vT, %du * PT);" % (b, LD, b))

    w("    for (int step = 0; step < ns; step++) {")
    if not XREG and not ROLL:
        for b in range(n):
            w("        x%d = %s;" % (b, cv("bld64(RX, v8, %du * P8)" % b)))
    if cfg.get("split", True):
        w("        for (int s = 0; s < nstage - 1; s++) {")
        stage(False)
        w("        }")
        w("        {")
        stage(True)
        w("        }")
    else:
        w("        for (int s = 0; s < nstage; s++) {")
        stage(None)
        w("        }")
    w("    }")
    w("    if (a.steps && a.steps_dec) a.steps[p] = rem - ns;")
    w("    if (a.status) {")
    w("        int st = 0;")
    w("        for (int r = 0; r < %d; r++) st |= isfinite(bld64(RX, v8, (uint32_t)r * P8)) ? 0 : 1;" % n)
    w("        a.status[p] |= st;")
    w("    }")
    w("}")
    w("#ifndef SFG_HOST_EMUL")
    w("hipError_t launch_%s(const EvolveArgs& a, hipStream_t s) {" % name)
    w("    hipLaunchKernelGGL(%s, dim3((unsigned)((a.P + %d) / %d)), dim3(%d), 0, s, a);" % (name, B - 1, B, B))
    w("    return hipGetLastError();")
    w("}")
    w("#endif")
    w("}  // namespace sfg")
    return "\n".join(o) + "\n"


def gen_kernel_sreg(L, prec, n, rows, cfg, vid):
    """Split register stepper ("sreg" mode, FP32 operator): S waves share one
    group of 64 parcels.  Every wave holds the whole stage input x (n FP32
    registers) and the RK increment sum of its own share of the rows (rows
    split into contiguous shares by term count); at the end of a stage each
    wave writes its rows' next-stage input to LDS and, after a workgroup
    barrier, every wave reads all n rows back.  The whole step loop is
    written once per share inside `if (wv == share)`, so a wave's registers
    hold only its own increment sums; both copies run the same barriers (the
    waves of a group step the same parcels, so their loop trip counts are
    equal).  The FP64 state is read per row chunk (x0) and written in the
    last stage by the row's owner.  Lanes past the batch stay alive for the
    barriers with a lane offset past every buffer (loads return 0, stores
    are dropped).
    """
    assert prec == "f32"
    T = "float"
    cv = lambda e: "(float)(%s)" % e
    name = "evolve_fast_L%d_%s_v%d" % (L, prec, vid)
    CHUNK, S = cfg["chunk"], cfg["split"]
    B = 64 * S
    used_p = sorted({p for (_, gs) in rows for (p, _) in gs})
    NPU = len(used_p)
    nq = min(n, 15)
    nblk = L // 2 + 1
    OMC, SLC = NPU, NPU + nq
    pidx = {q: i for i, q in enumerate(used_p)}
    nterm = [sum(len(tl) for (_, tl) in rows[r][1]) + 1 for r in range(n)]
    tot, acc_t, share = sum(nterm), 0, []
    for r in range(n):
        share.append(min(S - 1, int(S * (acc_t + nterm[r] / 2) / tot)))
        acc_t += nterm[r]
    o = []
    w = o.append
    w("// Generated by specfab_amd/tables/gen_fast.py -- do not edit.")
    w("// Split register stepper, L = %d (n = %d), FP32 stepping, FP64 state; variant %d: %s; row shares %s." %
      (L, n, vid, ", ".join("%s=%s" % kv for kv in sorted(cfg.items())),
       ", ".join("%d..%d" % (share.index(k), n - 1 - share[::-1].index(k)) for k in range(S))))
    w('#include "../sfg_kernels.h"')
    w("namespace sfg {")
    w("__global__ void __launch_bounds__(%d, %d) %s(EvolveArgs a) {" % (B, cfg["waves"], name))
    w("    const int64_t P = a.P;")
    w("    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;   // wv is uniform per wave")
    w("    const int64_t p = (int64_t)blockIdx.x * 64 + lane;")
    w("    const bool valid = p < P;")
    w("    const uint32_t P8 = (uint32_t)P * 8u;")
    w("    const uint32_t v8 = valid ? (uint32_t)p * 8u : 0x80000000u;")
    w("    const sfg_rsrc RX = make_rsrc(a.X, P8 * %du);" % n)
    w("    const int nstage = a.scheme ? 4 : 1;")
    w("    const int rem = valid ? (a.steps ? a.steps[p] : a.nsteps) : 0;")
    w("    const int ns = (a.steps && a.steps_dec) ? (rem < a.nsteps ? rem : a.nsteps) : rem;")
    w("    const double h = a.dt;")
    w("    // launch constants ([column][lane]: theta_p, omega_q, s_l) and the")
    w("    // stage input exchanged between the waves")
    w("    __shared__ float thl[%d][64];" % (NPU + nq + nblk))
    w("    __shared__ float xs[%d][64];" % n)
    w("    if (wv == 0 && valid) {")
    for q in used_p:
        w("        thl[%d][lane] = %s;" % (pidx[q], cv("a.theta[(int64_t)%d * P + p]" % q)))
    for q in range(nq):
        w("        thl[%d][lane] = %s;" % (OMC + q, cv("a.omega[(int64_t)%d * P + p]" % q)))
    for b in range(nblk):
        w("        thl[%d][lane] = %s;" % (SLC + b, cv("a.sl[(int64_t)%d * P + p]" % b)))
    w("    }")
    w("    float %s;" % ", ".join("x%d" % b for b in range(n)))
    for b in range(n):
        w("    x%d = %s;" % (b, cv("bld64(RX, v8, %du * P8)" % b)))
    w("    __syncthreads();")
    zc = [0]

    def zk():
        zc[0] += 1
        return zc[0]

    for sh in range(S):
        mine = [r for r in range(n) if share[r] == sh]
        chunks = [mine[c0:c0 + CHUNK] for c0 in range(0, len(mine), CHUNK)]
        w("    %sif (wv == %d) {" % ("" if sh == 0 else "} else ", sh))
        w("        float %s;" % ", ".join("y%d = 0.0f" % r for r in mine))
        w("        for (int step = 0; step < ns; step++) {")
        w("            for (int s = 0; s < nstage; s++) {")
        w("                const bool first = (s == 0), last = (s == nstage - 1);")
        w("                const float cy = %s;" % cv("(nstage == 1) ? h : ((s == 0 || s == 3) ? h / 6 : h / 3)"))
        w("                const float cx = %s;" % cv("(s == 2) ? h : h / 2"))
        w("                // a.zero is 0 at run time but unknown to the compiler: it keeps the")
        w("                // LDS reads of different row chunks and stages apart")
        w("                const int z = (int)a.zero * (step * 4 + s);")
        w("                float gD = 0.0f;")
        w("                if (a.use_ddrx) {")
        k = zk()
        terms = " + ".join("thl[%d][lane + z * %d] * x%d" % (OMC + q, k, q) for q in range(nq))
        w("                    gD = (%s) / x0;" % terms)
        w("                }")
        for chunk in chunks:
            k = zk()
            w("                {   // rows %s" % ", ".join(str(r) for r in chunk))
            w("                    const double %s;" % ", ".join("xo%d = bld64(RX, v8, %du * P8)" % (r, r) for r in chunk))
            w("                    float %s;" % ", ".join("acc%d = 0.0f" % r for r in chunk))
            byp = {}
            for r in chunk:
                for (pp, tl) in rows[r][1]:
                    byp.setdefault(pp, []).append((r, tl))
            for pp in sorted(byp):
                w("                    {")
                w("                        const float th = thl[%d][lane + z * %d];" % (pidx[pp], k))
                for (r, tl) in byp[pp]:
                    col, ti, v = tl[0]
                    w("                        float u%d = %s * x%d;" % (r, lit(v, prec), col))
                    for (col, ti, v) in tl[1:]:
                        w("                        u%d = fma(%s, x%d, u%d);" % (r, lit(v, prec), col, r))
                    w("                        acc%d = fma(th, u%d, acc%d);" % (r, r, r))
                w("                    }")
            for r in chunk:
                l = rows[r][0]
                w("                    {")
                w("                        const float f = fma(thl[%d][lane + z * %d] - gD, x%d, acc%d);"
                  % (SLC + l // 2, k, r, r))
                w("                        const float yv = fma(cy, f, first ? 0.0f : y%d);" % r)
                w("                        if (last) {")
                w("                            const double xn = xo%d + (double)yv;" % r)
                w("                            bst64(RX, v8, %du * P8, xn);" % r)
                w("                            xs[%d][lane] = (float)xn;" % r)
                w("                        } else {")
                w("                            y%d = yv;" % r)
                w("                            xs[%d][lane] = fma(cx, f, (float)xo%d);" % (r, r))
                w("                        }")
                w("                    }")
            w("                }")
        w("                __syncthreads();   // every wave's rows of the next stage input are in xs")
        w("                %s" % " ".join("x%d = xs[%d][lane];" % (b, b) for b in range(n)))
        w("                __syncthreads();   // read before the next stage overwrites xs")
        w("            }")
        w("        }")
    w("    }")
    w("    if (wv != 0 || !valid) return;")
    w("    if (a.steps && a.steps_dec) a.steps[p] = rem - ns;")
    w("    if (a.status) {")
    w("        int st = 0;")
    w("        for (int r = 0; r < %d; r++) st |= isfinite(bld64(RX, v8, (uint32_t)r * P8)) ? 0 : 1;" % n)
    w("        a.status[p] |= st;")
    w("    }")
    w("}")
    w("#ifndef SFG_HOST_EMUL")
    w("hipError_t launch_%s(const EvolveArgs& a, hipStream_t s) {" % name)
    w("    hipLaunchKernelGGL(%s, dim3((unsigned)((a.P + 63) / 64)), dim3(%d), 0, s, a);" % (name, B))
    w("    return hipGetLastError();")
    w("}")
    w("#endif")
    w("}  // namespace sfg")
    return "\n".join(o) + "\n"


def gen_kernel_dreg(L, prec, n, rows, cfg, vid):
    """Split resident stepper ("dreg" mode, FP64): S waves share one group of
    64 parcels, each on its own SIMD.  Every wave holds the whole stage input
    x (n FP64 values) and, for its own share of the rows (contiguous, split
    by term count), the RK increment sums y and the step-start state x0, so
    the state is read from HBM once per launch and written once at its end:
    no per-stage or per-step memory traffic besides LDS.  At the end of a
    stage each wave writes its rows' next-stage input (or, in the last
    stage, the new state) to LDS; after a workgroup barrier every wave reads
    all n rows back, and a second barrier keeps the next stage's writes
    behind those reads.  The step loop is written once per share inside
    `if (wv == share)`; the waves of a group step the same parcels, so both
    copies run the same number of barriers.  Lanes past the batch stay alive
    for the barriers with a lane offset past every buffer (loads return 0,
    stores are dropped).  The arithmetic of every row is that of the
    register steppers ("reg" mode), term for term.
    """
    assert prec == "f64"
    name = "evolve_fast_L%d_%s_v%d" % (L, prec, vid)
    CHUNK, S = cfg["chunk"], cfg["split"]
    B = 64 * S
    used_p = sorted({p for (_, gs) in rows for (p, _) in gs})
    NPU = len(used_p)
    nq = min(n, 15)
    nblk = L // 2 + 1
    OMC, SLC = NPU, NPU + nq
    pidx = {q: i for i, q in enumerate(used_p)}
    nterm = [sum(len(tl) for (_, tl) in rows[r][1]) + 1 for r in range(n)]
    tot, acc_t, share = sum(nterm), 0, []
    for r in range(n):
        share.append(min(S - 1, int(S * (acc_t + nterm[r] / 2) / tot)))
        acc_t += nterm[r]
    o = []
    w = o.append
    w("// Generated by specfab_amd/tables/gen_fast.py -- do not edit.")
    w("// Split resident stepper, L = %d (n = %d), FP64; variant %d: %s; row shares %s." %
      (L, n, vid, ", ".join("%s=%s" % kv for kv in sorted(cfg.items())),
       ", ".join("%d..%d" % (share.index(k), n - 1 - share[::-1].index(k)) for k in range(S))))
    w('#include "../sfg_kernels.h"')
    w("namespace sfg {")
    COLS = cfg.get("cols", False)
    w("__global__ void __launch_bounds__(%d, %d) %s(EvolveArgs a) {" % (B, 2 if COLS else 1, name))
    w("    const int64_t P = a.P;")
    w("    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;   // wv is uniform per wave")
    w("    const int64_t p = (int64_t)blockIdx.x * 64 + lane;")
    w("    const bool valid = p < P;")
    w("    const uint32_t P8 = (uint32_t)P * 8u;")
    w("    const uint32_t v8 = valid ? (uint32_t)p * 8u : 0x80000000u;")
    w("    const sfg_rsrc RX = make_rsrc(a.X, P8 * %du);" % n)
    w("    const int nstage = a.scheme ? 4 : 1;")
    w("    const int rem = valid ? (a.steps ? a.steps[p] : a.nsteps) : 0;")
    w("    const int ns = (a.steps && a.steps_dec) ? (rem < a.nsteps ? rem : a.nsteps) : rem;")
    w("    const double h = a.dt;")
    w("    // launch constants ([column][lane]: theta_p, omega_q, s_l) and the")
    w("    // stage input exchanged between the waves")
    w("    __shared__ double thl[%d][64];" % (NPU + nq + nblk))
    w("    __shared__ double xs[%d][64];" % n)
    if COLS:
        w("    __shared__ double ys[%d][64];   // RK increment sums" % n)
    cols = [("a.theta", q, pidx[q]) for q in used_p] + [("a.omega", q, OMC + q) for q in range(nq)] + \
           [("a.sl", b, SLC + b) for b in range(nblk)]
    for k in range(S):
        part = cols[k::S]
        w("    if (wv == %d && valid) {" % k)
        for (arr, q, c) in part:
            w("        thl[%d][lane] = %s[(int64_t)%d * P + p];" % (c, arr, q))
        w("    }")
    if not COLS:
        w("    double %s;" % ", ".join("x%d" % b for b in range(n)))
        for b in range(n):
            w("    x%d = bld64(RX, v8, %du * P8);" % (b, b))
        w("    __syncthreads();")
    zc = [0]

    def zk():
        zc[0] += 1
        return zc[0]

    for sh in range(S):
        mine = [r for r in range(n) if share[r] == sh]
        chunks = [mine[c0:c0 + CHUNK] for c0 in range(0, len(mine), CHUNK)]
        w("    %sif (wv == %d) {" % ("" if sh == 0 else "} else ", sh))
        if not COLS:
            w("        double %s;" % ", ".join("y%d = 0.0" % r for r in mine))
        if COLS:
            # only the columns this share's rows read (and its own rows) are
            # held; the stage input arrives through xs, also the first one
            # in the order of first use by the stage code (terms grouped by
            # parameter over the chunk's rows, then the rows' own diagonals),
            # so that the first groups can start while later columns arrive
            need = []
            for c0 in range(0, len(mine), CHUNK):
                byp0 = {}
                for r in mine[c0:c0 + CHUNK]:
                    for (pp, tl) in rows[r][1]:
                        byp0.setdefault(pp, []).append(tl)
                for pp in sorted(byp0):
                    for tl in byp0[pp]:
                        for (col, ti, v) in tl:
                            if col not in need:
                                need.append(col)
                for r in mine[c0:c0 + CHUNK]:
                    if r not in need:
                        need.append(r)
            w("        double %s;" % ", ".join("s%d = bld64(RX, v8, %du * P8)" % (r, r) for r in mine))
            w("        %s" % " ".join("xs[%d][lane] = s%d; ys[%d][lane] = 0.0;" % (r, r, r) for r in mine))
            w("        double %s;" % ", ".join("x%d" % b for b in need))
            w("        __syncthreads();")
            X = lambda q: ("x%d" % q) if q in need else ("xs[%d][lane]" % q)
        else:
            w("        double %s;" % ", ".join("s%d = x%d" % (r, r) for r in mine))
            X = lambda q: "x%d" % q
        w("        for (int step = 0; step < ns; step++) {")
        w("            for (int s = 0; s < nstage; s++) {")
        if COLS:
            w("                %s" % " ".join("x%d = xs[%d][lane];" % (b, b) for b in need))
        w("                const bool last = (s == nstage - 1);")
        w("                const double cy = (nstage == 1) ? h : ((s == 0 || s == 3) ? h / 6 : h / 3);")
        w("                const double cx = (s == 2) ? h : h / 2;")
        if COLS:
            w("                // stage kind as factors: the state takes the increment sum in the")
            w("                // last stage only, which also clears the sum and hands on the state")
            w("                const double lastf = last ? 1.0 : 0.0, keepf = last ? 0.0 : 1.0, cxe = last ? 0.0 : cx;")
        w("                // a.zero is 0 at run time but unknown to the compiler: it keeps the")
        w("                // LDS reads of different row chunks and stages apart")
        w("                const int z = (int)a.zero * (step * 4 + s);")
        w("                double gD = 0.0;")
        w("                if (a.use_ddrx) {")
        k = zk()
        terms = " + ".join("thl[%d][lane + z * %d] * %s" % (OMC + q, k, X(q)) for q in range(nq))
        w("                    gD = (%s) / %s;" % (terms, X(0)))
        w("                }")
        if COLS:
            w("                __syncthreads();   // every wave has read xs before any overwrites it")
        ngroup = 0
        for chunk in chunks:
            k = zk()
            w("                {   // rows %s" % ", ".join(str(r) for r in chunk))
            w("                    double %s;" % ", ".join("acc%d = 0.0" % r for r in chunk))
            byp = {}
            for r in chunk:
                for (pp, tl) in rows[r][1]:
                    byp.setdefault(pp, []).append((r, tl))
            for pp in sorted(byp):
                if ngroup % cfg["gb"] == 0:
                    # a basic-block boundary every gb (row chunk, parameter)
                    # groups: the code generator materialises a constant once
                    # per block for all its uses there, and over a whole stage
                    # those shared constants outgrow the scalar registers
                    w("                    if (z) asm volatile(\"\");")
                ngroup += 1
                w("                    {")
                w("                        const double th = thl[%d][lane + z * %d];" % (pidx[pp], k))
                for (r, tl) in byp[pp]:
                    col, ti, v = tl[0]
                    w("                        double u%d = %s * x%d;" % (r, lit(v, prec), col))
                    for (col, ti, v) in tl[1:]:
                        w("                        u%d = fma(%s, x%d, u%d);" % (r, lit(v, prec), col, r))
                    w("                        acc%d = fma(th, u%d, acc%d);" % (r, r, r))
                w("                    }")
            for r in chunk:
                l = rows[r][0]
                w("                    {")
                w("                        const double f = fma(thl[%d][lane + z * %d] - gD, x%d, acc%d);"
                  % (SLC + l // 2, k, r, r))
                # y is zero at every step's first stage (reset in the last one)
                if COLS:
                    # s + yv in the last stage (fma with 1: one rounding, as s + yv),
                    # s unchanged before it; the next stage input s + cx f before
                    # the last stage, the new state in it
                    w("                        const double yv = fma(cy, f, ys[%d][lane]);" % r)
                    w("                        s%d = fma(yv, lastf, s%d);" % (r, r))
                    w("                        ys[%d][lane] = yv * keepf;" % r)
                    w("                        xs[%d][lane] = fma(cxe, f, s%d);" % (r, r))
                else:
                    w("                        const double yv = fma(cy, f, y%d);" % r)
                    w("                        const double sn = s%d + yv, xn = fma(cx, f, s%d);" % (r, r))
                    w("                        y%d = last ? 0.0 : yv;" % r)
                    w("                        s%d = last ? sn : s%d;" % (r, r))
                    w("                        xs[%d][lane] = last ? sn : xn;" % r)
                w("                    }")
            w("                }")
        w("                __syncthreads();   // every wave's rows of the next stage input are in xs")
        if not COLS:
            w("                %s" % " ".join("x%d = xs[%d][lane];" % (b, b) for b in range(n)))
            w("                __syncthreads();   // read before the next stage overwrites xs")
        w("            }")
        w("        }")
        w("        %s" % " ".join("bst64(RX, v8, %du * P8, s%d);" % (r, r) for r in mine))
        w("        int st = 0;")
        w("        %s" % " ".join("st |= isfinite(s%d) ? 0 : 1;" % r for r in mine))
        w("        if (a.status && valid && st) atomicOr(a.status + p, st);")
    w("    }")
    w("    if (wv == 0 && valid && a.steps && a.steps_dec) a.steps[p] = rem - ns;")
    w("}")
    w("#ifndef SFG_HOST_EMUL")
    w("hipError_t launch_%s(const EvolveArgs& a, hipStream_t s) {" % name)
    w("    hipLaunchKernelGGL(%s, dim3((unsigned)((a.P + 63) / 64)), dim3(%d), 0, s, a);" % (name, B))
    w("    return hipGetLastError();")
    w("}")
    w("#endif")
    w("}  // namespace sfg")
    return "\n".join(o) + "\n"


def gen_kernel_lwin(L, prec, n, rows, cfg, vid):
    """LDS window stepper ("lwin" mode) for states too large for the
    registers of the waves of a parcel group (L >= 14).  S waves share 64
    parcels; the stage input of the group lives in LDS (n rows of the
    operator precision: FP64, or FP32 with the FP64 state), wave w owns a contiguous share of the rows (balanced by term
    count); their RK increment sums and next-stage inputs go through the YS
    and XA buffers (HBM / L2).  Per row chunk of its share (rows grouped greedily so that
    their coupled columns number at most cfg["window"]) a wave loads those
    columns of the stage input from LDS into registers and the chunk's
    step-start state rows from HBM (each row's owner is its only reader and
    writer), then evaluates the rows in sub-chunks of cfg["chunk"] with the
    coupling values as literals; the chunk's RK increment sums come from and
    go to the YS buffer (HBM / L2), prefetched with the state rows.  A stage
    ends with a workgroup barrier (every wave has read the stage input), the
    owners' copies of their next-stage rows from XA (the new state from X in
    the last stage) into LDS and a second barrier.
    The state is read once per row and stage and written in the last stage.  Lanes
    past the batch stay alive for the barriers with a lane offset past every
    buffer (loads return 0, stores are dropped).  The arithmetic of every row
    is that of the other steppers, term for term (FP32: the operator, the
    increment sums and the next-stage input in FP32, the state and its update
    x0 + sum in FP64).
    """
    F64 = prec == "f64"
    T = "double" if F64 else "float"
    LDT, STT = ("bld64", "bst64") if F64 else ("bld32", "bst32")
    cv = (lambda e: e) if F64 else (lambda e: "(float)(%s)" % e)
    zero = "0.0" if F64 else "0.0f"
    name = "evolve_fast_L%d_%s_v%d" % (L, prec, vid)
    WIN, CHUNK, S = cfg["window"], cfg["chunk"], cfg["split"]
    B = 64 * S
    used_p = sorted({p for (_, gs) in rows for (p, _) in gs})
    NPU = len(used_p)
    nq = min(n, 15)
    nblk = L // 2 + 1
    OMC, SLC = NPU, NPU + nq
    pidx = {q: i for i, q in enumerate(used_p)}
    need = [{col for (_, tl) in rows[r][1] for (col, _, _) in tl} | {r} for r in range(n)]
    nterm = [sum(len(tl) for (_, tl) in rows[r][1]) + 1 for r in range(n)]
    tot, acc_t, share = sum(nterm), 0, []
    for r in range(n):
        share.append(min(S - 1, int(S * (acc_t + nterm[r] / 2) / tot)))
        acc_t += nterm[r]
    chunks = []   # (share, rows, columns)
    for sh in range(S):
        cur, cols = [], set()
        for r in (r for r in range(n) if share[r] == sh):
            if cur and len(cols | need[r]) > WIN:
                chunks.append((sh, cur, sorted(cols)))
                cur, cols = [], set()
            cur.append(r)
            cols |= need[r]
        if cur:
            chunks.append((sh, cur, sorted(cols)))
    o = []
    w = o.append
    w("// Generated by specfab_amd/tables/gen_fast.py -- do not edit.")
    w("// LDS window stepper, L = %d (n = %d), %s; variant %d: %s; %d row chunks, %d column loads per stage." %
      (L, n, "FP64" if F64 else "FP32 stepping, FP64 state", vid, ", ".join("%s=%s" % kv for kv in sorted(cfg.items())), len(chunks),
       sum(len(c) for (_, _, c) in chunks)))
    w('#include "../sfg_kernels.h"')
    w("namespace sfg {")
    # four waves: one per SIMD; eight: two per SIMD (256 registers each)
    w("__global__ void __launch_bounds__(%d, %d) %s(EvolveArgs a) {" % (B, max(1, S // 4), name))
    w("    const int64_t P = a.P;")
    w("    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;   // wv is uniform per wave")
    w("    const int64_t p = (int64_t)blockIdx.x * 64 + lane;")
    w("    const bool valid = p < P;")
    w("    const uint32_t P8 = (uint32_t)P * 8u;")
    w("    const uint32_t v8 = valid ? (uint32_t)p * 8u : 0x80000000u;")
    w("    const sfg_rsrc RX = make_rsrc(a.X, P8 * %du);" % n)
    if F64:
        PT, vT = "P8", "v8"
    else:
        PT, vT = "P4", "v4"
        w("    const uint32_t P4 = (uint32_t)P * 4u;")
        w("    const uint32_t v4 = valid ? (uint32_t)p * 4u : 0x80000000u;")
    w("    const sfg_rsrc RYS = make_rsrc(a.YS, %s * %du);" % (PT, n))
    w("    const sfg_rsrc RXA = make_rsrc(a.XA, %s * %du);" % (PT, n))
    w("    const int nstage = a.scheme ? 4 : 1;")
    w("    const int rem = valid ? (a.steps ? a.steps[p] : a.nsteps) : 0;")
    w("    const int ns = (a.steps && a.steps_dec) ? (rem < a.nsteps ? rem : a.nsteps) : rem;")
    w("    const double h = a.dt;")
    w("    // launch constants ([column][lane]: theta_p, omega_q, s_l) and the")
    w("    // group's stage input")
    w("    __shared__ %s thl[%d][64];" % (T, NPU + nq + nblk))
    w("    __shared__ %s xs[%d][64];" % (T, n))
    consts = [("a.theta", q, pidx[q]) for q in used_p] + [("a.omega", q, OMC + q) for q in range(nq)] + \
             [("a.sl", b, SLC + b) for b in range(nblk)]
    for k in range(S):
        w("    if (wv == %d && valid) {" % k)
        for (arr, q, c) in consts[k::S]:
            w("        thl[%d][lane] = %s;" % (c, cv("%s[(int64_t)%d * P + p]" % (arr, q))))
        w("    }")
    zc = [0]

    def zk():
        zc[0] += 1
        return zc[0]

    for sh in range(S):
        mine = [r for r in range(n) if share[r] == sh]
        w("    %sif (wv == %d) {" % ("" if sh == 0 else "} else ", sh))
        w("        %s" % " ".join("xs[%d][lane] = %s;" % (r, cv("bld64(RX, v8, %du * P8)" % r)) for r in mine))
        w("        __syncthreads();")
        w("        for (int step = 0; step < ns; step++) {")
        w("            for (int s = 0; s < nstage; s++) {")
        w("                const bool first = (s == 0), last = (s == nstage - 1);")
        w("                const %s cy = %s;" % (T, cv("(nstage == 1) ? h : ((s == 0 || s == 3) ? h / 6 : h / 3)")))
        w("                const %s cx = %s;" % (T, cv("(s == 2) ? h : h / 2")))
        w("                // a.zero is 0 at run time but unknown to the compiler: it keeps the")
        w("                // LDS reads of different row chunks and stages apart")
        w("                const int z = (int)a.zero * (step * 4 + s);")
        w("                // row offsets (r + zu) P8 vary with the stage for the compiler, so")
        w("                // they are formed where used instead of held in scalar registers")
        w("                const uint32_t zu = (uint32_t)z;")
        w("                %s gD = %s;" % (T, zero))
        w("                if (a.use_ddrx) {")
        k = zk()
        terms = " + ".join("thl[%d][lane + z * %d] * xs[%d][lane]" % (OMC + q, k, q) for q in range(nq))
        w("                    gD = (%s) / xs[0][lane];" % terms)
        w("                }")
        ngroup = 0
        for (csh, crow, ccols) in chunks:
            if csh != sh:
                continue
            kc = zk()
            w("                {   // rows %d..%d: %d columns" % (crow[0], crow[-1], len(ccols)))
            # the z term keeps the compiler from reusing one chunk's column
            # values in another (which would keep them all live)
            w("                    const %s %s;" % (T, ", ".join("c%d = xs[%d][lane + z * %d]" % (c, c, kc) for c in ccols)))
            for c0 in range(0, len(crow), CHUNK):
                sub = crow[c0:c0 + CHUNK]
                k = zk()
                w("                    {")
                w("                        const double %s;" % ", ".join("xo%d = bld64(RX, v8, (%du + zu) * P8)" % (r, r)
                                                                     for r in sub))
                w("                        const %s %s;" % (T, ", ".join(
                    "yo%d = first ? %s : %s(RYS, %s, (%du + zu) * %s)" % (r, zero, LDT, vT, r, PT) for r in sub)))
                w("                        %s %s;" % (T, ", ".join("acc%d = %s" % (r, zero) for r in sub)))
                byp = {}
                for r in sub:
                    for (pp, tl) in rows[r][1]:
                        byp.setdefault(pp, []).append((r, tl))
                for pp in sorted(byp):
                    if ngroup % cfg["gb"] == 0:
                        # a basic-block boundary every gb (row chunk, parameter)
                        # groups keeps the constants shared within a block
                        # inside the scalar registers
                        w("                        if (z) asm volatile(\"\");")
                    ngroup += 1
                    w("                        {")
                    w("                            const %s th = thl[%d][lane + z * %d];" % (T, pidx[pp], k))
                    for (r, tl) in byp[pp]:
                        col, ti, v = tl[0]
                        w("                            %s u%d = %s * c%d;" % (T, r, lit(v, prec), col))
                        for (col, ti, v) in tl[1:]:
                            w("                            u%d = fma(%s, c%d, u%d);" % (r, lit(v, prec), col, r))
                        w("                            acc%d = fma(th, u%d, acc%d);" % (r, r, r))
                    w("                        }")
                for r in sub:
                    l = rows[r][0]
                    w("                        {")
                    w("                            const %s f = fma(thl[%d][lane + z * %d] - gD, c%d, acc%d);"
                      % (T, SLC + l // 2, k, r, r))
                    w("                            const %s yv = fma(cy, f, yo%d);" % (T, r))
                    w("                            if (last) {")
                    w("                                bst64(RX, v8, (%du + zu) * P8, xo%d + %s);"
                      % (r, r, "yv" if F64 else "(double)yv"))
                    w("                            } else {")
                    w("                                %s(RXA, %s, (%du + zu) * %s, fma(cx, f, %s));"
                      % (STT, vT, r, PT, ("xo%d" % r) if F64 else ("(float)xo%d" % r)))
                    w("                                %s(RYS, %s, (%du + zu) * %s, yv);" % (STT, vT, r, PT))
                    w("                            }")
                    w("                        }")
                w("                    }")
            w("                }")
        w("                __syncthreads();   // every wave has read this stage's input")
        if F64:
            w("                {")
            w("                    const sfg_rsrc RN = last ? RX : RXA;   // this lane's own stores, read back")
            w("                    %s" % " ".join("xs[%d][lane] = bld64(RN, v8, (%du + zu) * P8);" % (r, r) for r in mine))
            w("                }")
        else:
            w("                if (last) {   // this lane's own stores, read back")
            w("                    %s" % " ".join("xs[%d][lane] = (float)bld64(RX, v8, (%du + zu) * P8);" % (r, r)
                                                for r in mine))
            w("                } else {")
            w("                    %s" % " ".join("xs[%d][lane] = bld32(RXA, v4, (%du + zu) * P4);" % (r, r) for r in mine))
            w("                }")
        w("                __syncthreads();   // the next stage input is complete")
        w("            }")
        w("        }")
        w("        int st = 0;")
        if F64:
            w("        %s" % " ".join("st |= isfinite(xs[%d][lane]) ? 0 : 1;" % r for r in mine))
        else:
            w("        %s" % " ".join("st |= isfinite(bld64(RX, v8, %du * P8)) ? 0 : 1;" % r for r in mine))
        w("        if (a.status && valid && st) atomicOr(a.status + p, st);")
    w("    }")
    w("    if (wv == 0 && valid && a.steps && a.steps_dec) a.steps[p] = rem - ns;")
    w("}")
    w("#ifndef SFG_HOST_EMUL")
    w("hipError_t launch_%s(const EvolveArgs& a, hipStream_t s) {" % name)
    w("    hipLaunchKernelGGL(%s, dim3((unsigned)((a.P + 63) / 64)), dim3(%d), 0, s, a);" % (name, B))
    w("    return hipGetLastError();")
    w("}")
    w("#endif")
    w("}  // namespace sfg")
    return "\n".join(o) + "\n"


def gen_kernel_win(L, prec, n, rows, cfg, vid):
    """Window stepper ("win" mode) for states too large for a lane's
    registers (L >= 14 FP64, L >= 18 FP32).

    Rows are taken in order and grouped greedily into chunks whose coupled
    columns (the union of the columns their terms read, plus the rows
    themselves) number at most cfg["window"]; per chunk those columns of the
    stage input are loaded into registers once (the operator is banded:
    |dl| <= 4, |dm| <= 4), then the chunk's rows are evaluated in sub-chunks
    of cfg["chunk"] rows with the coupling values as literals.  Stage buffers
    (YS increment sum, XA / XB stage inputs) go through HBM; launch constants
    are staged in LDS.  Stage input / output: RK4 X -> XA -> XB -> XA -> X;
    Euler X -> XB (FP64) -> X.

    With cfg["split"] = S > 1 a workgroup is S waves on the same 64 parcels;
    wave w evaluates the w-th share of the row chunks and a workgroup
    barrier closes every stage (the next stage's columns come from all
    shares).  Lanes past the batch stay alive for the barriers with a lane
    offset past every buffer (0x80000000; the buffers are < 2^31 bytes), so
    their loads return 0 and their stores are dropped.
    """
    T = "double" if prec == "f64" else "float"
    LDT = "bld64" if prec == "f64" else "bld32"
    STT = "bst64" if prec == "f64" else "bst32"
    W = 8 if prec == "f64" else 4
    cv = (lambda e: e) if prec == "f64" else (lambda e: "(float)(%s)" % e)
    name = "evolve_fast_L%d_%s_v%d" % (L, prec, vid)
    WIN, CHUNK, B = cfg["window"], cfg["chunk"], cfg["block"]
    SPLIT = cfg.get("split", 1)
    assert B == 64 * SPLIT or SPLIT == 1
    NL = 64 if SPLIT > 1 else B       # parcels per workgroup
    used_p = sorted({p for (_, gs) in rows for (p, _) in gs})
    NPU = len(used_p)
    nq = min(n, 15)
    nblk = L // 2 + 1
    OMC, SLC = NPU, NPU + nq
    pidx = {q: i for i, q in enumerate(used_p)}
    need = [{col for (_, tl) in rows[r][1] for (col, _, _) in tl} | {r} for r in range(n)]
    chunks, cur, cols = [], [], set()
    for r in range(n):
        if cur and len(cols | need[r]) > WIN:
            chunks.append((cur, sorted(cols)))
            cur, cols = [], set()
        cur.append(r)
        cols |= need[r]
    chunks.append((cur, sorted(cols)))
    # shares of the chunks per wave, balanced by term count
    nterm = [sum(len(tl) for (_, tl) in rows[r][1]) + 1 for r in range(n)]
    cw = [sum(nterm[r] for r in crow) for (crow, _) in chunks]
    tot, acc, share = sum(cw), 0, []
    for c in cw:
        share.append(min(SPLIT - 1, int(SPLIT * (acc + c / 2) / tot)))
        acc += c
    o = []
    w = o.append
    w("// Generated by specfab_amd/tables/gen_fast.py -- do not edit.")
    w("// Window stepper, L = %d (n = %d), %s; variant %d: %s; %d row chunks, %d column loads per stage." %
      (L, n, "FP64" if prec == "f64" else "FP32 stepping, FP64 state", vid,
       ", ".join("%s=%s" % kv for kv in sorted(cfg.items())), len(chunks), sum(len(c) for _, c in chunks)))
    w('#include "../sfg_kernels.h"')
    w("namespace sfg {")
    w("__global__ void __launch_bounds__(%d, %d) %s(EvolveArgs a) {" % (B, cfg["waves"], name))
    w("    const int64_t P = a.P;")
    if SPLIT > 1:
        w("    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;   // wv is uniform per wave")
        w("    const int64_t p = (int64_t)blockIdx.x * 64 + lane;")
        w("    const bool valid = p < P;")
        w("    const uint32_t P8 = (uint32_t)P * 8u, PT = (uint32_t)P * %du;" % W)
        w("    const uint32_t v8 = valid ? (uint32_t)p * 8u : 0x80000000u;")
        w("    const uint32_t vT = valid ? (uint32_t)p * %du : 0x80000000u;" % W)
    else:
        w("    const int lane = threadIdx.x;")
        w("    const int64_t p = (int64_t)blockIdx.x * %d + threadIdx.x;" % B)
        w("    if (p >= P) return;")
        w("    const uint32_t P8 = (uint32_t)P * 8u, PT = (uint32_t)P * %du;" % W)
        w("    const uint32_t v8 = (uint32_t)p * 8u, vT = (uint32_t)p * %du;" % W)
    w("    const sfg_rsrc RX = make_rsrc(a.X, P8 * %du);" % n)
    w("    const sfg_rsrc RYS = make_rsrc(a.YS, PT * %du);" % n)
    w("    const sfg_rsrc RXA = make_rsrc(a.XA, PT * %du);" % n)
    w("    const sfg_rsrc RXB = make_rsrc(a.XB, PT * %du);" % n)
    w("    const sfg_rsrc RE = make_rsrc(a.XB, P8 * %du);   // Euler output (FP64 rows)" % n)
    w("    const int nstage = a.scheme ? 4 : 1;")
    w("    const bool euler = (nstage == 1);")
    if SPLIT > 1:
        w("    const int rem = valid ? (a.steps ? a.steps[p] : a.nsteps) : 0;")
    else:
        w("    const int rem = a.steps ? a.steps[p] : a.nsteps;")
    w("    const int ns = (a.steps && a.steps_dec) ? (rem < a.nsteps ? rem : a.nsteps) : rem;")
    w("    const double h = a.dt;")
    w("    __shared__ %s thl[%d][%d];" % (T, NPU + nq + nblk, NL))
    ind = "    "
    if SPLIT > 1:
        w("    if (wv == 0 && valid) {")
        ind = "        "
    for q in used_p:
        w("%sthl[%d][lane] = %s;" % (ind, pidx[q], cv("a.theta[(int64_t)%d * P + p]" % q)))
    for q in range(nq):
        w("%sthl[%d][lane] = %s;" % (ind, OMC + q, cv("a.omega[(int64_t)%d * P + p]" % q)))
    for b in range(nblk):
        w("%sthl[%d][lane] = %s;" % (ind, SLC + b, cv("a.sl[(int64_t)%d * P + p]" % b)))
    if SPLIT > 1:
        w("    }")
        w("    __syncthreads();")
    w("    for (int step = 0; step < ns; step++) {")
    w("        for (int s = 0; s < nstage; s++) {")
    w("            const bool first = (s == 0), last = (s == nstage - 1);")
    w("            const %s cy = %s;" % (T, cv("euler ? h : ((s == 0 || s == 3) ? h / 6 : h / 3)")))
    w("            const %s cx = %s;" % (T, cv("(s == 2) ? h : h / 2")))
    w("            // stage input (s > 0) and output buffers")
    w("            const sfg_rsrc RIN = (s == 2) ? RXB : RXA;")
    w("            const sfg_rsrc ROUT = (s == 1) ? RXB : RXA;")
    w("            // a.zero is 0 at run time but unknown to the compiler: it keeps the")
    w("            // LDS reads of different chunks and stages apart")
    w("            const int z = (int)a.zero * (step * 4 + s);")
    zc = [0]

    def zk():
        zc[0] += 1
        return zc[0]

    def loads(colset, indent):
        # column values of the stage input into registers
        w("%s%s %s;" % (indent, T, ", ".join("c%d" % c for c in colset)))
        w("%sif (first) {" % indent)
        for c in colset:
            w("%s    c%d = %s;" % (indent, c, cv("bld64(RX, v8, %du * P8)" % c)))
        w("%s} else {" % indent)
        for c in colset:
            w("%s    c%d = %s(RIN, vT, %du * PT);" % (indent, c, LDT, c))
        w("%s}" % indent)

    # <D>: omega . x / x0 over the l <= 4 rows
    w("            %s gD = (%s)0;" % (T, T))
    w("            if (a.use_ddrx) {")
    loads(list(range(nq)), "                ")
    k = zk()
    terms = " + ".join("thl[%d][lane + z * %d] * c%d" % (OMC + q, k, q) for q in range(nq))
    w("                gD = (%s) / c0;" % terms)
    w("            }")
    for ci, (crow, ccols) in enumerate(chunks):
        if SPLIT > 1 and (ci == 0 or share[ci] != share[ci - 1]):
            w("            %sif (wv == %d) {" % ("" if ci == 0 else "} else ", share[ci]))
        w("            {   // rows %d..%d: %d columns" % (crow[0], crow[-1], len(ccols)))
        loads(ccols, "                ")
        for c0 in range(0, len(crow), CHUNK):
            sub = crow[c0:c0 + CHUNK]
            k = zk()
            w("                {")
            w("                    %s %s;" % (T, ", ".join("acc%d = (%s)0" % (r, T) for r in sub)))
            byp = {}
            for r in sub:
                for (pp, tl) in rows[r][1]:
                    byp.setdefault(pp, []).append((r, tl))
            for pp in sorted(byp):
                w("                    {")
                w("                        const %s th = thl[%d][lane + z * %d];" % (T, pidx[pp], k))
                for (r, tl) in byp[pp]:
                    col, ti, v = tl[0]
                    w("                        %s u%d = %s * c%d;" % (T, r, lit(v, prec), col))
                    for (col, ti, v) in tl[1:]:
                        w("                        u%d = fma(%s, c%d, u%d);" % (r, lit(v, prec), col, r))
                    w("                        acc%d = fma(th, u%d, acc%d);" % (r, r, r))
                w("                    }")
            for r in sub:
                l = rows[r][0]
                w("                    {")
                w("                        const %s f = fma(thl[%d][lane + z * %d] - gD, c%d, acc%d);"
                  % (T, SLC + l // 2, k, r, r))
                w("                        const double x0a = bld64(RX, v8, %du * P8);" % r)
                w("                        if (euler) {")
                w("                            bst64(RE, v8, %du * P8, x0a + (double)(cy * f));" % r)
                w("                        } else {")
                w("                            const %s ys = first ? cy * f : fma(cy, f, %s(RYS, vT, %du * PT));"
                  % (T, LDT, r))
                w("                            if (last) {")
                w("                                bst64(RX, v8, %du * P8, x0a + (double)ys);" % r)
                w("                            } else {")
                w("                                %s(RYS, vT, %du * PT, ys);" % (STT, r))
                w("                                %s(ROUT, vT, %du * PT, fma(cx, f, (%s)x0a));" % (STT, r, T))
                w("                            }")
                w("                        }")
                w("                    }")
            w("                }")
        w("            }")
    if SPLIT > 1:
        w("            }")
        w("            __syncthreads();   // the next stage reads every wave's rows")
    w("        }")
    w("        if (euler) {")
    if SPLIT > 1:
        for wvi in range(SPLIT):
            rs = [r for ci, (crow, _) in enumerate(chunks) if share[ci] == wvi for r in crow]
            if rs:
                w("            if (wv == %d) for (int r = %d; r <= %d; r++) bst64(RX, v8, (uint32_t)r * P8, bld64(RE, v8, (uint32_t)r * P8));" % (wvi, rs[0], rs[-1]))
        w("            __syncthreads();")
    else:
        w("            for (int r = 0; r < %d; r++) bst64(RX, v8, (uint32_t)r * P8, bld64(RE, v8, (uint32_t)r * P8));" % n)
    w("        }")
    w("    }")
    if SPLIT > 1:
        w("    if (wv != 0 || !valid) return;")
    w("    if (a.steps && a.steps_dec) a.steps[p] = rem - ns;")
    w("    if (a.status) {")
    w("        int st = 0;")
    w("        for (int r = 0; r < %d; r++) st |= isfinite(bld64(RX, v8, (uint32_t)r * P8)) ? 0 : 1;" % n)
    w("        a.status[p] |= st;")
    w("    }")
    w("}")
    w("#ifndef SFG_HOST_EMUL")
    w("hipError_t launch_%s(const EvolveArgs& a, hipStream_t s) {" % name)
    w("    hipLaunchKernelGGL(%s, dim3((unsigned)((a.P + %d) / %d)), dim3(%d), 0, s, a);" % (name, NL - 1, NL, B))
    w("    return hipGetLastError();")
    w("}")
    w("#endif")
    w("}  // namespace sfg")
    return "\n".join(o) + "\n"


def scratch_mask(cfg):
    """stage buffers (n x P each) a variant reads or writes: 1 = YS (increment
    sum), 2 = XA, 4 = XB (stage inputs); see the make_rsrc lines of each
    generator"""
    mode = cfg["mode"]
    if mode == "lit":
        return 1 | 2
    if mode == "reg":
        return 2 if cfg["xn"] == "mem" else 0
    if mode in ("sreg", "dreg"):
        return 0
    if mode == "lwin":
        return 1 | 2
    return 1 | 2 | 4   # win


def gen_dispatch(kinds):
    """kinds: list of (L, prec, variant configs)."""
    o = ["// Generated by specfab_amd/tables/gen_fast.py -- do not edit.",
         '#include "../sfg_fast.h"', "namespace sfg {"]
    for (L, prec, cfgs) in kinds:
        for v in range(len(cfgs)):
            o.append("hipError_t launch_evolve_fast_L%d_%s_v%d(const EvolveArgs& a, hipStream_t s);" % (L, prec, v))
    o.append("int fast_variants(int L, int prec) {")
    for (L, prec, cfgs) in kinds:
        o.append("    if (L == %d && prec == %d) return %d;" % (L, 0 if prec == "f64" else 1, len(cfgs)))
    o.append("    return 0;")
    o.append("}")
    o.append("#ifndef SFG_HOST_EMUL")
    for (L, prec, cfgs) in kinds:
        for v in range(len(cfgs)):
            o.append("__global__ void evolve_fast_L%d_%s_v%d(EvolveArgs a);" % (L, prec, v))
    o.append("const void* fast_kernel(int L, int prec, int variant, int* block, int* parcels) {")
    for (L, prec, cfgs) in kinds:
        for v, cfg in enumerate(cfgs):
            ppb = 64 if cfg.get("split", 1) > 1 else cfg["block"]
            o.append("    if (L == %d && prec == %d && variant == %d) { *block = %d; *parcels = %d; "
                     "return (const void*)&evolve_fast_L%d_%s_v%d; }"
                     % (L, 0 if prec == "f64" else 1, v, cfg["block"], ppb, L, prec, v))
    o.append("    return nullptr;")
    o.append("}")
    o.append("#endif")
    o.append("int fast_buffers(int L, int prec, int variant) {")
    for (L, prec, cfgs) in kinds:
        for v, cfg in enumerate(cfgs):
            o.append("    if (L == %d && prec == %d && variant == %d) return %d;"
                     % (L, 0 if prec == "f64" else 1, v, scratch_mask(cfg)))
    o.append("    return BUF_YS | BUF_XA | BUF_XB;")
    o.append("}")
    o.append("hipError_t launch_evolve_fast(int L, int prec, int variant, const EvolveArgs& a, hipStream_t s) {")
    o.append("    if (a.P == 0) return hipSuccess;")
    for (L, prec, cfgs) in kinds:
        for v in range(len(cfgs)):
            o.append("    if (L == %d && prec == %d && variant == %d) return launch_evolve_fast_L%d_%s_v%d(a, s);"
                     % (L, 0 if prec == "f64" else 1, v, L, prec, v))
    o.append("    return hipErrorInvalidValue;")
    o.append("}")
    o.append("}  // namespace sfg")
    return "\n".join(o) + "\n"


def gen_moment_maps(exe):
    """Mandel a2 / a4 of the l <= 4 block of a state as literal FMAs over the
    maps' nonzero coefficients (in increasing row order, so the sums equal
    those of a loop over all 15 rows)."""
    a2, a4 = {}, {}
    for line in subprocess.check_output([exe, "maps"]).decode().split("\n"):
        if not line:
            continue
        tag, i, q, v = line.split()
        (a2 if tag == "A2" else a4).setdefault(int(i), []).append((int(q), float.fromhex(v)))
    o = ["// Generated by specfab_amd/tables/gen_fast.py -- do not edit.",
         "// Mandel a2 and a4 of the l <= 4 block of a real-basis state x (xq[q] = row q):",
         "// a2v[i] = a2s[i] / x0 + (i < 3 ? 1/3 : 0), a4v[i][j] = a4s[i][j] / x0, with the",
         "// nonzero coefficients of the maps of src/moments.f90:164-178 as literals.",
         "#pragma once", "namespace sfg {",
         "__device__ __forceinline__ void moment_maps(const double (&xq)[15], double (&a2s)[6], double (&a4s)[6][6]) {"]
    for i in range(6):
        terms = sorted(a2.get(i, []))
        o.append("    { double s = 0.0;" + "".join(" s = fma(%s, xq[%d], s);" % (v.hex(), q) for q, v in terms)
                 + " a2s[%d] = s; }" % i)
    for i in range(6):
        for j in range(i, 6):
            terms = sorted(a4.get(6 * i + j, []))
            assert terms == sorted(a4.get(6 * j + i, [])), "a4 map not symmetric"
            o.append("    { double s = 0.0;" + "".join(" s = fma(%s, xq[%d], s);" % (v.hex(), q) for q, v in terms)
                     + " a4s[%d][%d] = s;%s }" % (i, j, "" if i == j else " a4s[%d][%d] = s;" % (j, i)))
    o.append("}")
    o.append("}  // namespace sfg")
    return "\n".join(o) + "\n"


def write_if_changed(path, text):
    if os.path.exists(path) and open(path).read() == text:
        return
    with open(path + ".tmp", "w") as f:
        f.write(text)
    os.replace(path + ".tmp", path)


def main():
    outdir = sys.argv[1]
    os.makedirs(outdir, exist_ok=True)
    exe = os.path.join(outdir, "dump_basis")
    srcs = [os.path.join(HERE, "dump_basis.cpp"), os.path.join(CSRC, "sfg_basis.cpp")]
    if not os.path.exists(exe) or any(os.path.getmtime(s) > os.path.getmtime(exe) for s in srcs):
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", CSRC, "-I", outdir] + srcs + ["-o", exe])
    kinds = []
    for prec, Ls in FAST_L.items():
        for L in Ls:
            n, rows = dump(L, exe)
            kind = "wide" if L in WIDE_L[prec] else ("tuned" if L in TUNED_L else "default")
            vs = VARIANTS[(prec, kind)]
            for vid, cfg in enumerate(vs):
                gen = {"reg": gen_kernel_reg, "win": gen_kernel_win, "sreg": gen_kernel_sreg,
                       "dreg": gen_kernel_dreg, "lwin": gen_kernel_lwin}.get(cfg["mode"], gen_kernel)
                write_if_changed(os.path.join(outdir, "sfg_fast_L%d_%s_v%d.hip" % (L, prec, vid)),
                                 gen(L, prec, n, rows, cfg, vid))
            kinds.append((L, prec, vs))
    # drop generated kernels of an older variant list
    keep = {"sfg_fast_L%d_%s_v%d.hip" % (L, prec, v) for (L, prec, cfgs) in kinds for v in range(len(cfgs))}
    for f in os.listdir(outdir):
        if f.startswith("sfg_fast_L") and f.endswith(".hip") and f not in keep:
            os.remove(os.path.join(outdir, f))
    write_if_changed(os.path.join(outdir, "sfg_fast_dispatch.hip"), gen_dispatch(kinds))
    write_if_changed(os.path.join(outdir, "sfg_moment_maps.h"), gen_moment_maps(exe))
    print("generated %d fast kernels" % sum(len(c) for (_, _, c) in kinds))


if __name__ == "__main__":
    main()
