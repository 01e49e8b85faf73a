"""CPU oracle of the learned-proposal MLP.  TEST INFRASTRUCTURE ONLY.

FP64 numpy forward/backward of nin -> nh -> nh -> 2*nu (ReLU) with the
Gaussian negative log-likelihood, rounding to BF16 (round to nearest even)
exactly where the device kernels do: GEMM inputs (features, weights,
activations, upstream gradients); biases and accumulation stay wide.  Any
remaining difference to the device comes from FP32 accumulation order.
The MLP has no counterpart in the reference; its loss/gradient formulas are
pinned in tests by central finite differences in pure FP64.
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    """Round float values to BF16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, nin, nh, npad=32):
    o = 0
    out = []
    for shape in ((nin, nh), (nh,), (nh, nh), (nh,), (nh, npad), (npad,)):
        n = int(np.prod(shape))
        out.append(np.asarray(params[o:o + n], dtype=np.float64).reshape(shape))
        o += n
    return out


def loss_and_grad(params, X, U, nin, nh, nu, emulate_bf16=True, npad=32):
    """Mean Gaussian NLL over the batch and its gradient (flat, params layout)."""
    W1, b1, W2, b2, W3, b3 = unpack(params, nin, nh, npad)
    r = bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    B = X.shape[0]
    Xb = r(X)
    Z1 = Xb @ r(W1) + b1
    A1 = r(np.maximum(Z1, 0))
    Z2 = A1 @ r(W2) + b2
    A2 = r(np.maximum(Z2, 0))
    O = A2 @ r(W3) + b3
    mu = O[:, :nu]
    sr = O[:, nu:2 * nu]
    s = np.clip(sr, -7, 3)
    iv = np.exp(-2 * s)
    e = U - mu
    L = np.sum(0.5 * e * e * iv + s, axis=1)
    dO = np.zeros_like(O)
    dO[:, :nu] = -e * iv / B
    dO[:, nu:2 * nu] = np.where((sr > -7) & (sr < 3), 1 - e * e * iv, 0.0) / B
    dOb = r(dO)
    gW3 = A2.T @ dOb
    gb3 = dOb.sum(0)
    dZ2 = r((dOb @ r(W3).T) * (Z2 > 0))
    gW2 = A1.T @ dZ2
    gb2 = dZ2.sum(0)
    dZ1 = r((dZ2 @ r(W2).T) * (Z1 > 0))
    gW1 = Xb.T @ dZ1
    gb1 = dZ1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return float(L.mean()), g


def forward(params, X, nin, nh, nu, npad=32):
    W1, b1, W2, b2, W3, b3 = unpack(params, nin, nh, npad)
    A1 = np.maximum(X @ W1 + b1, 0)
    A2 = np.maximum(A1 @ W2 + b2, 0)
    return (A2 @ W3 + b3)[:, :2 * nu]
