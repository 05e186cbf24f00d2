"""CPU FP64 oracle for the warm-start MLP -- TEST INFRASTRUCTURE ONLY.

Restates, in numpy float64, the network the GPU kernels train:
x -> relu(x W1 + b1) -> relu(. W2 + b2) -> . W3 + b3, with inputs and targets
standardised, loss = mean((pred - target)^2) over batch and outputs.  The
gradients are the textbook back-propagation of that loss; they are pinned by
a central finite-difference check on a tiny network (tests/test_mlp_oracle.py).
There is no reference counterpart: the reference has no learned component.
"""
from __future__ import annotations

import numpy as np


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, n_in, hid, n_out):
    p = np.asarray(params, np.float64)
    o = 0
    out = []
    for shape in [(n_in, hid), (hid,), (hid, hid), (hid,), (hid, n_out), (n_out,)]:
        size = int(np.prod(shape))
        out.append(p[o:o + size].reshape(shape))
        o += size
    return out


def forward_backward(params, X, T, n_in, hid, n_out, in_mean, in_std, out_mean, out_std, bf16=False):
    """Loss and flat gradient.  With bf16=True the GEMM operands are rounded to
    bfloat16 exactly where the GPU kernels round them (inputs, weights,
    hidden activations, output gradient, hidden gradients)."""
    W1, b1, W2, b2, W3, b3 = unpack(params, n_in, hid, n_out)
    def r(a):
        return bf16_round(np.asarray(a, np.float32)) if bf16 else np.asarray(a, np.float64)

    x = (np.asarray(X, np.float64) - in_mean) / in_std
    t = (np.asarray(T, np.float64) - out_mean) / out_std
    xb = r(x)
    z1 = xb @ r(W1) + b1
    h1 = np.maximum(z1, 0.0)
    z2 = r(h1) @ r(W2) + b2
    h2 = np.maximum(z2, 0.0)
    p = r(h2) @ r(W3) + b3
    B = x.shape[0]
    loss = np.mean((p - t) ** 2)
    dp = 2.0 * (p - t) / (B * n_out)
    dpb = r(dp)
    gW3 = r(h2).T @ dpb
    gb3 = dp.sum(axis=0)
    dz2 = (dpb @ r(W3).T) * (h2 > 0)
    dz2b = r(dz2)
    gW2 = r(h1).T @ dz2b
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2b @ r(W2).T) * (h1 > 0)
    dz1b = r(dz1)
    gW1 = xb.T @ dz1b
    gb1 = dz1.sum(axis=0)
    grad = np.concatenate([g.ravel() for g in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, grad, p * out_std + out_mean
