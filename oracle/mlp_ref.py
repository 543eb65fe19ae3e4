"""CPU restatement of the log-likelihood surrogate network (numpy, float64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
network's forward pass, loss and gradient, never as the product path.

The network has no counterpart in the reference (the reference samples with
the exact likelihood only).  It is an MLP nin -> H -> H -> H -> 1 with GELU
activations (the erf form), laid out as the C ABI's parameter vector
W1[H x nin], b1[H], W2[H x H], b2, W3[H x H], b3, W4[1 x H], b4[1] at the
offsets pofd_mlp_layout reports.  With ``emulate=True`` the bfloat16
roundings of the GPU path are applied where it applies them (standardised
inputs, weights, hidden activations); the arithmetic between them is
float64.  The gradient is that of the mean squared error on standardised
targets, with the roundings treated as the identity (the derivative the
GPU's backward pass computes).

Pinned by central finite differences of the loss with ``emulate=False``
(tests/test_oracle_mlp.py), which any dropped term, transposed operand or
wrong activation derivative fails.
"""
import numpy as np
from scipy.special import erf

_SQRT1_2 = 0.70710678118654752
_INV_SQRT_2PI = 0.39894228040143268


def bf16(a):
    """Round float32 values to bfloat16 (nearest, ties to even); returned
    as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    r = np.where(nan, (u >> 16) | 0x40, r).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z * _SQRT1_2))


def gelu_prime(z):
    return 0.5 * (1.0 + erf(z * _SQRT1_2)) + z * _INV_SQRT_2PI * np.exp(
        -0.5 * z * z)


def unpack(w, offsets, nin, H):
    """The eight parameter blocks of the flat vector."""
    w = np.asarray(w, dtype=np.float64)
    o = offsets
    return [w[o[0]:o[0] + H * nin].reshape(H, nin), w[o[1]:o[1] + H],
            w[o[2]:o[2] + H * H].reshape(H, H), w[o[3]:o[3] + H],
            w[o[4]:o[4] + H * H].reshape(H, H), w[o[5]:o[5] + H],
            w[o[6]:o[6] + H].reshape(1, H), w[o[7]:o[7] + 1]]


def standardize(x, xmean, xstd):
    """(x - mean) / std as the GPU forms it: float32 arithmetic with the
    reciprocal of std."""
    x = np.asarray(x, dtype=np.float32)
    istd = (np.float32(1.0) / np.asarray(xstd, dtype=np.float32))
    return ((x - np.asarray(xmean, dtype=np.float32)) * istd).astype(
        np.float64)


def whiten(x, xmean, wmat):
    """W (x - mean) for each row: the difference in float32 as the GPU
    forms it, the product in float64 (the GPU accumulates it in float32)."""
    d = (np.asarray(x, dtype=np.float32) -
         np.asarray(xmean, dtype=np.float32)).astype(np.float64)
    return d @ np.asarray(wmat, dtype=np.float32).astype(np.float64).T


def forward(w, offsets, nin, H, xs, emulate=True):
    """Standardised prediction of standardised inputs xs[n x nin] and the
    intermediates the backward pass needs."""
    rnd = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(w, offsets, nin, H)
    Ws = [rnd(W1), rnd(W2), rnd(W3), rnd(W4)]
    bs = [b1, b2, b3, b4]
    a = rnd(xs)
    acts, zs = [a], []
    for l in range(3):
        z = a @ Ws[l].T + bs[l]
        zs.append(z)
        a = rnd(gelu(z))
        acts.append(a)
    out = (a @ Ws[3].T + bs[3])[:, 0]
    return out, dict(W=Ws, z=zs, a=acts)


def loss_grad(w, offsets, nin, H, xs, ys, emulate=True):
    """Mean squared error of standardised targets ys and its gradient with
    respect to the flat parameter vector."""
    out, c = forward(w, offsets, nin, H, xs, emulate)
    M = xs.shape[0]
    r = out - ys
    loss = np.mean(r * r)
    g = np.zeros(len(w))
    o = offsets
    d = 2.0 * r / M                                  # dL/dout
    g[o[6]:o[6] + H] = d @ c["a"][3]                 # W4
    g[o[7]] = d.sum()                                # b4
    dA = np.outer(d, c["W"][3][0])                   # dL/dA3
    for l in (2, 1, 0):
        dZ = dA * gelu_prime(c["z"][l])
        K = nin if l == 0 else H
        g[o[2 * l]:o[2 * l] + H * K] = (dZ.T @ c["a"][l]).ravel()
        g[o[2 * l + 1]:o[2 * l + 1] + H] = dZ.sum(axis=0)
        dA = dZ @ c["W"][l]
    return loss, g
