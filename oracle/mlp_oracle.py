"""numpy restatement of the learned-closure network and its training loss --
TEST INFRASTRUCTURE ONLY (the checker of the HIP training kernels).

Parameters are the flat FP32 master vector of the library (W1 [32][10] at 0, b1 at
320, W2 [32][32] at 352, b2 at 1376, W3 [32] at 1408, b3 at 1440); inputs are the
sample table rows 0-9 normalised as (raw - shift) * scale; the loss of a sample is
c^2 P + 2 c Q + R with c = clamp(MLP(x), -e^-1, e^-1), P, Q, R from rows 10-12.
Everything here is float64.  Pinned by central finite differences of `loss` and by
the torch BF16 conversion (tests/test_closure_oracle.py).
"""
from __future__ import annotations

import numpy as np

OFF = dict(W1=0, b1=320, W2=352, b2=1376, W3=1408, b3=1440)
NPARAM = 1441
CMAX = float(np.exp(-1.0))       # output clamp: |c| <= e^-1
CMIN = -CMAX
LOG2E, LN2 = 1.4426950408889634, 0.6931471805599453


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[0:320].reshape(32, 10), p[320:352], p[352:1376].reshape(32, 32), p[1376:1408], p[1408:1440],
            p[1440])


def silu(a):
    return a / (1.0 + np.exp(-a))


def dsilu(a):
    s = 1.0 / (1.0 + np.exp(-a))
    return s * (1.0 + a * (1.0 - s))


def forward(p, x):
    """y for normalised inputs x (n, 10), exact arithmetic (float64)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    a1 = x @ W1.T + b1
    h1 = silu(a1)
    a2 = h1 @ W2.T + b2
    return silu(a2) @ W3 + b3


def bf16(a):
    """Nearest BF16 value (ties to even) of float32(a), as float64."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(p, shift, scale, raw):
    """Inference as the GPU rounds it: the inputs, W1, W2, W3 and the first hidden layer
    in BF16, and the hidden layers in base-2 units u = -log2(e) a (scale, b1, b2 times
    -log2(e) and W3 times -ln 2, each rounded to FP32; a unit's output u / (1 + 2^u) is
    -SiLU(a) / ln 2); exact arithmetic otherwise."""
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, np.float32).astype(np.float64))
    f32 = lambda v: np.asarray(v, np.float64).astype(np.float32)
    scale2 = f32(-LOG2E * np.asarray(scale, np.float32).astype(np.float64))
    x = (np.asarray(raw, np.float32) - np.asarray(shift, np.float32)) * scale2
    x = bf16(x)
    b1u = f32(-LOG2E * b1).astype(np.float64)
    b2u = f32(-LOG2E * b2).astype(np.float64)
    w3u = f32(-LN2 * bf16(W3)).astype(np.float64)
    u1 = x @ bf16(W1).T + b1u
    h1 = bf16(u1 / (1.0 + np.exp2(u1)))
    u2 = h1 @ bf16(W2).T + b2u
    return (u2 / (1.0 + np.exp2(u2))) @ w3u + b3


def normalise(table, shift, scale, rows):
    raw = np.asarray(table, np.float64)[:10, rows].T
    return (raw - np.asarray(shift, np.float64)) * np.asarray(scale, np.float64)


def loss(p, x, P, Q, R, inv_norm=1.0):
    c = np.clip(forward(p, x), CMIN, CMAX)
    return float(np.sum(c * c * P + 2 * c * Q + R) * inv_norm)


def loss_grad(p, x, P, Q, R, inv_norm=1.0):
    """(loss, gradient over the flat parameter vector) by reverse-mode differentiation."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    a1 = x @ W1.T + b1
    h1 = silu(a1)
    a2 = h1 @ W2.T + b2
    h2 = silu(a2)
    y = h2 @ W3 + b3
    clamped = (y > CMAX) | (y < CMIN)
    c = np.clip(y, CMIN, CMAX)
    L = float(np.sum(c * c * P + 2 * c * Q + R) * inv_norm)
    dy = np.where(clamped, 0.0, (2 * c * P + 2 * Q) * inv_norm)
    g = np.zeros(NPARAM)
    g[1408:1440] = h2.T @ dy
    g[1440] = dy.sum()
    d2 = (dy[:, None] * W3[None, :]) * dsilu(a2)
    g[352:1376] = (d2.T @ h1).ravel()
    g[1376:1408] = d2.sum(0)
    d1 = (d2 @ W2) * dsilu(a1)
    g[0:320] = (d1.T @ x).ravel()
    g[320:352] = d1.sum(0)
    return L, g
