"""Pins the CPU restatement of the surrogate network (oracle/mlp_ref.py):
its backpropagated gradients against central finite differences of its own
loss (exact FP64 mode), and the bf16 rounding helper against exact cases."""
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
