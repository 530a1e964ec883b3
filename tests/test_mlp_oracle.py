"""Pins the CPU restatement of the surrogate network (oracle/mlp_ref.py):
its backpropagated gradients against central finite differences of its own
loss (exact FP64 mode), the bf16 rounding helper against exact cases and
the table GELU of the fused inference kernel against the exact one."""
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


def test_gelu_table_error_bound():
    """|gelu_table - z Phi(z)| <= 4e-6 on [-5, 5] (linear interpolation of Phi
    with h = 1/102.4: h^2/8 max|Phi''| = 3e-6, times |z|), <= 3e-7 |z| + 4e-6
    beyond; exact at the
    nodes; GELU(z) - GELU(-z) = z up to the interpolation error."""
    from scipy.special import erf
    z = np.linspace(-5, 5, 200001)
    exact = 0.5 * z * (1 + erf(z / np.sqrt(2)))
    assert np.abs(M.gelu_table(z) - exact).max() <= 4e-6
    zb = np.linspace(-200, 200, 20001)
    exb = 0.5 * zb * (1 + erf(zb / np.sqrt(2)))
    assert np.all(np.abs(M.gelu_table(zb) - exb) <= 3e-7 * np.abs(zb) + 4e-6)
    zn = (np.arange(0, 1024) - 512.0) / float(np.float32(102.4))
    assert np.abs(M.gelu_table(zn) - 0.5 * zn * (1 + erf(zn / np.sqrt(2)))).max() < 1e-15
    assert np.abs(M.gelu_table(z) - M.gelu_table(-z) - z).max() <= 6e-6


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
