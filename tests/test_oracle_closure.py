"""The closure network's CPU oracle pinned without reference to its own
formulas: the convolution against a direct sum over pixels and taps, and the
gradient against central finite differences of the loss."""
import numpy as np

import closure_oracle as CO


def test_param_count():
    # 1 -> 32: 9*32 + 32; three 32 -> 32: 3 (9*32*32 + 32); 32 -> 1: 9*32 + 1
    assert CO.param_count() == 320 + 3 * 9248 + 289 == 28353


def test_convolution_matches_direct_sum():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 8, 8, 3))
    w = rng.standard_normal((4, 9, 3))
    b = rng.standard_normal(4)
    z = CO.conv(x, w, b)
    for (bb, yy, xx, co) in [(0, 0, 0, 0), (1, 7, 3, 2), (0, 4, 7, 3), (1, 0, 7, 1)]:
        s = b[co]
        for ky in range(3):
            for kx in range(3):
                s += x[bb, (yy + ky - 1) % 8, (xx + kx - 1) % 8, :] @ w[co, 3 * ky + kx, :]
        assert abs(z[bb, yy, xx, co] - s) < 1e-12


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = rng.standard_normal(CO.param_count()) * 0.2
    x = rng.standard_normal((2, 8, 8))
    t = rng.standard_normal((2, 8, 8))
    _, g = CO.loss_and_grad(p, x, t)
    idx = rng.choice(p.size, size=40, replace=False)
    idx = np.concatenate([idx, [0, 288, 288 + 32, p.size - 1, p.size - 2]])  # layer edges, biases
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (CO.loss_and_grad(p + e, x, t)[0] - CO.loss_and_grad(p - e, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
