"""Pins the CPU restatement of the surrogate network (oracle/mlp_ref.py):
its backpropagated gradients against central finite differences of its own
loss (exact FP64 mode), the bf16 rounding helper against exact cases and
the polynomial GELU of the fused inference kernel against the exact one."""
import numpy as np

from oracle import mlp_ref as M


def _setup(seed=3, B=5):
    rng = np.random.default_rng(seed)
    n = sum(M.DIMS[i] * M.DIMS[i + 1] + M.DIMS[i + 1] for i in range(4))
    p = rng.normal(0, 0.08, n)
    X = rng.normal(0, 1, (B, 36))
    T = rng.normal(0, 1, (B, 9))
    norm = np.concatenate([rng.normal(0, 0.1, 36), rng.uniform(0.5, 2, 36), rng.normal(0, 0.1, 9),
                           rng.uniform(0.5, 2, 9)])
    return p, X, T, norm


def test_bf16_rounding():
    assert M.bf16(1.0) == 1.0
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties to even -> 1
    assert M.bf16(1.0 + 2.0 ** -8) == 1.0
    assert M.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert M.bf16(-2.5) == -2.5
    x = np.random.default_rng(0).normal(size=1000)
    assert np.max(np.abs(M.bf16(x) - x) / np.abs(x)) <= 2.0 ** -8


def test_gelu_poly_error_bound():
    """|gelu_poly - z Phi(z)| <= 3.7e-5 on [-15, 15] (the bound quoted in the
    kernel), 1.1e-5 |z| + 4e-5 beyond; exact symmetry GELU(z) - GELU(-z) = z;
    the value at 0 and the saturation beyond the clamp."""
    from scipy.special import erf
    z = np.linspace(-15, 15, 300001)
    exact = 0.5 * z * (1 + erf(z / np.sqrt(2)))
    assert np.abs(M.gelu_poly(z) - exact).max() <= 3.7e-5
    zb = np.linspace(-200, 200, 40001)
    exb = 0.5 * zb * (1 + erf(zb / np.sqrt(2)))
    assert np.all(np.abs(M.gelu_poly(zb) - exb) <= 1.1e-5 * np.abs(zb) + 4e-5)
    assert np.allclose(M.gelu_poly(z) - M.gelu_poly(-z), z, rtol=0, atol=1e-12)
    assert M.gelu_poly(0.0) == 0.0
    # beyond the clamp GELU(z) / z is the constant Phi_poly(4.25) ~ 1
    assert abs(M.gelu_poly(100.0) / 100.0 - 1.0) < 1.1e-5
    assert abs(M.gelu_poly(100.0) / 100.0 - M.gelu_poly(4.25) / 4.25) < 1e-15


def test_gradients_match_finite_differences():
    p, X, T, norm = _setup()
    loss, g = M.loss_and_grad(p, X, T, norm, bf=False)
    rng = np.random.default_rng(11)
    n = p.size
    # one index in every parameter block plus random ones
    sizes = [M.DIMS[i] * M.DIMS[i + 1] + M.DIMS[i + 1] for i in range(4)]
    starts = np.cumsum([0] + sizes[:-1])
    idx = list(starts) + list(starts + np.array(sizes) - 1) + list(rng.integers(0, n, 12))
    h = 1e-6
    for k in idx:
        e = np.zeros(n)
        e[k] = h
        lp = M.loss_and_grad(p + e, X, T, norm, bf=False)[0]
        lm = M.loss_and_grad(p - e, X, T, norm, bf=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 * max(1.0, abs(g[k])) + 1e-9, (k, fd, g[k])
