"""CPU reference of the warm-start network (test infrastructure only).

Forward and backward pass of in -> H -> H -> out with GELU and the mean
squared error, in float64.  With bf16=True the operands of every matrix
product are rounded to bfloat16 (round to nearest even) at the same points as
the device computation (inputs, weights, hidden activations, output and
hidden-layer gradients), so that the remaining difference from the device is
the FP32 accumulation order; with bf16=False it is the exact float64
computation, which the finite-difference test pins.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf


def bf16(a) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def split(params, dims):
    """Parameter vector (W1, b1, W2, b2, W3, b3; W row-major in x out) -> list."""
    i, h, o = dims
    shapes = [(i, h), (h,), (h, h), (h,), (h, o), (o,)]
    out, off = [], 0
    for s in shapes:
        k = int(np.prod(s))
        out.append(np.asarray(params[off:off + k], dtype=np.float64).reshape(s))
        off += k
    assert off == len(params)
    return out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def loss_grad(params, dims, X, T, bf16_products: bool = False):
    """(loss, gradient vector) of sum (Y - T)^2 / (rows * out)."""
    r = bf16 if bf16_products else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = split(params, dims)
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    rows, out = T.shape
    xb = r(X)
    z1 = xb @ r(W1) + b1
    h1 = r(gelu(z1))
    z2 = h1 @ r(W2) + b2
    h2 = r(gelu(z2))
    y = h2 @ r(W3) + b3
    d = y - T
    loss = float((d * d).sum() / (rows * out))
    dy = 2.0 * d / (rows * out)
    dyb = r(dy)
    gW3, gb3 = h2.T @ dyb, dy.sum(axis=0)
    dz2 = (dyb @ r(W3).T) * gelu_grad(z2)
    dz2b = r(dz2)
    gW2, gb2 = h1.T @ dz2b, dz2.sum(axis=0)
    dz1 = (dz2b @ r(W2).T) * gelu_grad(z1)
    gW1, gb1 = xb.T @ r(dz1), dz1.sum(axis=0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return loss, g


def predict(params, dims, X, bf16_products: bool = False):
    r = bf16 if bf16_products else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3 = split(params, dims)
    h1 = r(gelu(r(X) @ r(W1) + b1))
    h2 = r(gelu(h1 @ r(W2) + b2))
    return h2 @ r(W3) + b3
