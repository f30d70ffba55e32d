"""FP64 restatement of the warm-start MLP (TEST INFRASTRUCTURE: used by tests/ only).

Same network as accelgraph_amd/csrc/ag_mlp.hip: standardised 5 inputs (padded to 16) -> 64 ->
64 -> 1 with ReLU, mean-squared loss on the standardised target. `forward_backward(...,
bf16=True)` applies the same BF16 roundings the GPU applies to MFMA operands (weights, inputs,
first hidden activations, the two backpropagated deltas) and computes everything else in FP64;
with bf16=False it is the exact FP64 gradient, which tests check against finite differences.
"""
from __future__ import annotations

import numpy as np

IN, H, NF = 16, 64, 5
P_W1, P_B1 = 0, IN * H
P_W2 = P_B1 + H
P_B2 = P_W2 + H * H
P_W3 = P_B2 + H
P_B3 = P_W3 + H
P_COUNT = P_B3 + 1


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[P_W1:P_B1].reshape(IN, H), p[P_B1:P_W2], p[P_W2:P_B2].reshape(H, H),
            p[P_B2:P_W3], p[P_W3:P_B3], p[P_B3])


def standardise(X, y, in_mean, in_std, y_mean, y_std):
    X = np.asarray(X, np.float32)
    xs = np.zeros((X.shape[0], IN), np.float32)
    xs[:, :NF] = (X[:, :NF] - np.asarray(in_mean, np.float32)[:NF]) / np.asarray(in_std, np.float32)[:NF]
    ys = None if y is None else (np.asarray(y, np.float32) - np.float32(y_mean)) / np.float32(y_std)
    return xs, ys


def forward_backward(p, X, y, in_mean, in_std, y_mean, y_std, bf16_ops=True):
    """Returns (loss, grad[P_COUNT], prediction in standardised units)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    xs, ys = standardise(X, y, in_mean, in_std, y_mean, y_std)
    r = bf16 if bf16_ops else (lambda v: np.asarray(v, np.float64))
    x = r(xs)
    W1r, W2r = (bf16(W1.astype(np.float32)), bf16(W2.astype(np.float32))) if bf16_ops else (W1, W2)
    z1 = x @ W1r + b1
    h1 = np.maximum(z1, 0.0)
    h1r = r(h1.astype(np.float32)) if bf16_ops else h1
    z2 = h1r @ W2r + b2
    h2 = np.maximum(z2, 0.0)
    yhat = h2 @ w3 + b3
    B = X.shape[0]
    d = yhat - ys.astype(np.float64)
    loss = float(np.mean(d * d))
    dy = 2.0 * d / B
    g = np.zeros(P_COUNT)
    g[P_W3:P_B3] = h2.T @ dy
    g[P_B3] = dy.sum()
    dz2 = np.outer(dy, w3) * (h2 > 0)
    g[P_B2:P_W3] = dz2.sum(0)
    dz2r = r(dz2.astype(np.float32)) if bf16_ops else dz2
    g[P_W2:P_B2] = (h1r.T @ dz2r).ravel()
    dh1 = dz2r @ W2r.T
    dz1 = dh1 * (h1 > 0)
    g[P_B1:P_W2] = dz1.sum(0)
    dz1r = r(dz1.astype(np.float32)) if bf16_ops else dz1
    g[P_W1:P_B1] = (x.T @ dz1r).ravel()
    return loss, g, yhat
