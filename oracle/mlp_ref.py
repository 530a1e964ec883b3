"""CPU restatement of the surrogate network (test infrastructure only).

Forward pass, MSE loss and backpropagated gradients of the network of
specfab_amd (36 -> 256 -> 256 -> 256 -> 9, GELU, include/specfab_gpu.h) in
FP64 numpy.  With `bf16=True` the operands are rounded to bfloat16 (round to
nearest even) at the points where the GPU kernels round them (inputs,
weights, activations, output gradients), so the remaining difference to the
GPU is FP32 accumulation order; with `bf16=False` the math is exact FP64 and
is pinned by central finite differences in tests/test_mlp_oracle.py.

The network has no counterpart in the reference; its reference is the plain
homogenization (oracle/sf_oracle.c, checked against the reference).
"""
import math

import numpy as np

DIMS = (36, 256, 256, 256, 9)


def bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    """Flat parameters -> [(W (in, out), b (out))] * 4."""
    p = np.asarray(params, dtype=np.float64)
    out, off = [], 0
    for i in range(4):
        n_in, n_out = DIMS[i], DIMS[i + 1]
        W = p[off:off + n_in * n_out].reshape(n_in, n_out)
        off += n_in * n_out
        b = p[off:off + n_out]
        off += n_out
        out.append((W, b))
    assert off == p.size
    return out


def gelu(z):
    return 0.5 * z * (1 + np.vectorize(math.erf)(z / math.sqrt(2)))


def gelu_table(z):
    """GELU of the fused inference kernel (specfab_amd/csrc/sfg_surrogate.hip,
    gelu_tab): Phi linearly interpolated between nodes z_k = (k - 512) / 102.4,
    k = 0..1024, index clamped to [0, 1024), in FP64."""
    z = np.asarray(z, dtype=np.float64)
    sc = float(np.float32(102.4))
    t = np.clip(z * sc + 512.0, 0.0, float.fromhex("0x1.fffffep+9"))
    k = np.floor(t)
    zk = (k - 512.0) / sc
    zk1 = (k + 1 - 512.0) / sc
    p0 = 0.5 * (1 + np.vectorize(math.erf)(zk / math.sqrt(2)))
    p1 = 0.5 * (1 + np.vectorize(math.erf)(zk1 / math.sqrt(2)))
    return z * (p0 + (p1 - p0) * (t - k))


def dgelu(z):
    return 0.5 * (1 + np.vectorize(math.erf)(z / math.sqrt(2))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


def forward(params, X, norm, bf=True, act=None):
    """Normalised outputs (B, 9) and the saved intermediates; `act` replaces
    the exact GELU (e.g. gelu_table)."""
    act = act or gelu
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    im, isd = norm[0:36], norm[36:72]
    layers = unpack(params)
    xin = r((np.asarray(X, np.float64) - im) / isd)
    h, zs, acts = xin, [], [xin]
    for i, (W, b) in enumerate(layers):
        z = h @ r(W) + b
        if i < 3:
            zs.append(z)
            h = r(act(z))
            acts.append(h)
        else:
            y = z
    return y, zs, acts


def loss_and_grad(params, X, T, norm, bf=True):
    """MSE over the normalised outputs and its gradient (flat, like params)."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    om, osd = norm[72:81], norm[81:90]
    y, zs, acts = forward(params, X, norm, bf)
    tn = (np.asarray(T, np.float64) - om) / osd
    B = X.shape[0]
    d = y - tn
    loss = float(np.mean(d * d))
    dy = 2.0 * d / (B * 9)
    layers = unpack(params)
    grads = [None] * 4
    dnext_b, dnext = r(dy), dy
    for i in range(3, -1, -1):
        W, b = layers[i]
        gW = acts[i].T @ dnext_b
        gb = dnext.sum(0)
        grads[i] = (gW, gb)
        if i == 0:
            break
        da = dnext_b @ r(W).T
        dz = da * dgelu(zs[i - 1])
        dnext_b, dnext = r(dz), dz
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat
