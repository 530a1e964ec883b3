"""sfg_evolve_noalloc (the stepping entry point that allocates nothing: its
workspace is a static device array) against sfg_evolve on the same inputs,
through the C-ABI: two calls in a row, the second with fewer parcels (the
array keeps the first call's contents), FP64 and FP32C, bit-identical
results (the same kernels); device memory in use unchanged by the calls
(nothing allocated); a batch beyond the array and a truncation whose tables
are not on the device yet are refused with a message."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _desc(L, P, x, G, g0, nsteps, prec):
    from specfab_amd import _lib
    d = _lib.EvolveDesc()
    d.L, d.P = L, P
    d.procs = _lib.SFG_LROT | _lib.SFG_REG | _lib.SFG_DDRX
    d.scheme, d.precision, d.nsteps, d.dt = _lib.SFG_RK4, prec, nsteps, 0.5
    d.iota, d.zeta = 1.0, 0.0
    d.ugrad_dev, d.gam0_dev, d.x_dev = G.data_ptr(), g0.data_ptr(), x.data_ptr()
    return d


def _field(P, L, seed):
    import math
    import torch
    g = torch.Generator(device="cuda:0")
    g.manual_seed(seed)
    n = (L + 1) * (L + 2) // 2
    G = torch.randn((P, 3, 3), generator=g, device="cuda:0", dtype=torch.float64) * 1e-3
    G -= (G.diagonal(dim1=1, dim2=2).sum(-1) / 3)[:, None, None] * torch.eye(3, device="cuda:0", dtype=torch.float64)
    g0 = 10.0 ** (torch.rand(P, generator=g, device="cuda:0", dtype=torch.float64) * 2 - 4)
    x = torch.randn((n, P), generator=g, device="cuda:0", dtype=torch.float64) * 1e-3
    x[0] = 1 / math.sqrt(4 * math.pi)
    return G.contiguous(), g0.contiguous(), x.contiguous()


def test_noalloc_matches_evolve():
    import torch
    from specfab_amd import _lib
    lib = _lib.lib()
    ctx = ctypes.c_void_p()
    _lib.check(lib.sfg_ctx_create(0, ctypes.byref(ctx)))
    try:
        L = 12
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        # a fresh context has no tables on the device: refused, not uploaded
        G, g0, x = _field(64, L, 1)
        d = _desc(L, 64, x, G, g0, 2, _lib.SFG_FP64)
        rc = lib.sfg_evolve_noalloc(ctx, ctypes.byref(d), stream)
        assert rc == -1   # SFG_ERR_ARG
        assert b"not on the device" in lib.sfg_last_error()
        # tables of L = 12 onto the device (P = 0 allocates no workspace)
        d0 = _desc(L, 0, x, G, g0, 2, _lib.SFG_FP64)
        _lib.check(lib.sfg_evolve(ctx, ctypes.byref(d0), stream))
        # the first launch of each kernel in the process loads its code object
        # (the runtime's one-time allocation): one call per precision first
        for prec in (_lib.SFG_FP64, _lib.SFG_FP32C):
            G, g0, x = _field(128, L, 3)
            _lib.check(lib.sfg_evolve_noalloc(ctx, ctypes.byref(_desc(L, 128, x, G, g0, 1, prec)), stream))
        torch.cuda.synchronize()
        for prec in (_lib.SFG_FP64, _lib.SFG_FP32C):
            # a fresh context runs variant 0 (stage buffers in the workspace:
            # 1,816 B per parcel at L = 12, so up to ~73,000 parcels)
            for P, seed, steps in ((60013, 7, 5), (4099, 8, 3)):
                G, g0, x = _field(P, L, seed)
                xa, xb = x.clone(), x.clone()
                torch.cuda.synchronize()
                free0 = torch.cuda.mem_get_info()[0]
                _lib.check(lib.sfg_evolve_noalloc(ctx, ctypes.byref(_desc(L, P, xa, G, g0, steps, prec)), stream))
                torch.cuda.synchronize()
                assert torch.cuda.mem_get_info()[0] == free0
                _lib.check(lib.sfg_evolve(ctx, ctypes.byref(_desc(L, P, xb, G, g0, steps, prec)), stream))
                torch.cuda.synchronize()
                assert torch.equal(xa, xb), (prec, P)
                assert not torch.equal(xa, x)
                print("noalloc prec=%d P=%d steps=%d: bit-identical to sfg_evolve" % (prec, P, steps), flush=True)
        # beyond the static workspace: refused with the sizes in the message
        P = 1 << 20
        G, g0, x = _field(P, L, 9)
        rc = lib.sfg_evolve_noalloc(ctx, ctypes.byref(_desc(L, P, x, G, g0, 1, _lib.SFG_FP64)), stream)
        assert rc != 0 and b"bytes of workspace" in lib.sfg_last_error()
    finally:
        lib.sfg_ctx_destroy(ctx)
