"""Pins the MLP oracle: the analytic loss gradient against central finite
differences (pure FP64, no BF16 emulation)."""
import numpy as np

from oracle import mlp as OM


def test_gradient_finite_differences():
    rng = np.random.default_rng(0)
    nin, nh, nu, B = 8, 32, 4, 16
    n = nin * nh + nh + nh * nh + nh + nh * 32 + 32
    p = rng.standard_normal(n) * 0.3
    X = rng.standard_normal((B, nin))
    U = rng.random((B, nu))
    L, g = OM.loss_and_grad(p, X, U, nin, nh, nu, emulate_bf16=False)
    idx = rng.choice(n, 60, replace=False)
    # only output columns < 2*nu carry gradient; padded columns are exactly zero
    for i in idx:
        e = np.zeros(n)
        h = 1e-6
        e[i] = h
        Lp, _ = OM.loss_and_grad(p + e, X, U, nin, nh, nu, emulate_bf16=False)
        Lm, _ = OM.loss_and_grad(p - e, X, U, nin, nh, nu, emulate_bf16=False)
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * (1 + abs(fd)), (i, fd, g[i])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5])
    np.testing.assert_array_equal(OM.bf16(x), [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5])
