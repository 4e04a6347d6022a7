"""CPU: the network oracle's hand-written backward pass against central finite differences, and
the BF16 rounding helper against exactly representable values."""
import numpy as np

import oracle.mlp as om


def test_bf16_round():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 3.0, -2.5, 1.0 + 3 * 2 ** -9], dtype=np.float32)
    r = om.bf16_round(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: nearest even is 1; 1 + 3*2^-9 rounds up to 1 + 2^-7
    assert list(r) == [1.0, 1.0, 1.0 + 2 ** -7, 3.0, -2.5, 1.0 + 2 ** -7]


def test_backward_vs_finite_differences():
    rng = np.random.default_rng(3)
    dims = (5, 7, 4)
    n = sum(int(np.prod(s)) for s in om.shapes(*dims))
    p = rng.normal(size=n) * 0.5
    X = rng.normal(size=(6, 5))
    T = rng.normal(size=(6, 4))
    _, g = om.loss_and_grad(p, X, T, dims)
    h = 1e-6
    fd = np.zeros(n)
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        fd[i] = (om.loss_and_grad(p + e, X, T, dims)[0] - om.loss_and_grad(p - e, X, T, dims)[0]) / (2 * h)
    assert np.abs(fd - g).max() < 1e-7 * max(1.0, np.abs(g).max())
    # a dropped activation derivative would fail
    assert np.abs(g).max() > 1e-3
