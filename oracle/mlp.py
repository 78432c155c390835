"""CPU restatement of the closure network (28 -> 128 -> 128 -> 25, GELU).

TEST INFRASTRUCTURE ONLY: the checker of the device network's forward pass,
gradient and Adam step; never measured or shipped.  Only tests/ may use it.

The network has no counterpart in the reference: it is specified by the
closure's own interface (include/vipp_gravity.h, VIPP_MLP_*).  In FP64, with
the option to round weights and activations to FP16 exactly where the device
forward pass does (numpy's float16 conversion rounds to nearest even, as the
device's does), so the two evaluate the same function.
"""
from __future__ import annotations

import math

import numpy as np

IN, H, OUT = 28, 128, 25
OW1 = 0
OB1 = OW1 + H * IN
OW2 = OB1 + H
OB2 = OW2 + H * H
OW3 = OB2 + H
OB3 = OW3 + OUT * H
NPAR = OB3 + OUT

_erf = np.vectorize(math.erf, otypes=[np.float64])


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OW1:OB1].reshape(H, IN), p[OB1:OW2], p[OW2:OB2].reshape(H, H), p[OB2:OW3], p[OW3:OB3].reshape(OUT, H), p[OB3:NPAR])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


def f16(a):
    return np.asarray(a, dtype=np.float64).astype(np.float16).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(params, x, half=True):
    """Outputs and the activations the backward pass needs."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    r = f16 if half else (lambda a: np.asarray(a, dtype=np.float64))
    x0 = r(x)
    W1, W2, W3 = r(W1), r(W2), r(W3)
    z1 = x0 @ W1.T + b1
    a1 = r(gelu(z1))
    z2 = a1 @ W2.T + b2
    a2 = r(gelu(z2))
    y = a2 @ W3.T + b3
    return y, (x0, z1, a1, z2, a2, W1, W2, W3)


def loss_grad(params, x, t, half=True, w=None):
    """Mean squared error over all n x 25 outputs, each sample weighted by w
    (sum_j w_j |y_j - t_j|^2 / (25 sum_j w_j)), and its gradient (FP16
    roundings, when emulated, are passed straight through)."""
    y, (x0, z1, a1, z2, a2, W1, W2, W3) = forward(params, x, half)
    n = len(x)
    w = np.ones(n) if w is None else np.asarray(w, dtype=np.float64)
    d = y - np.asarray(t, dtype=np.float64)
    loss = float((w[:, None] * d * d).sum() / (w.sum() * OUT))
    dy = 2.0 * w[:, None] * d / (w.sum() * OUT)
    gW3, gb3 = dy.T @ a2, dy.sum(0)
    dz2 = (dy @ W3) * gelu_grad(z2)
    gW2, gb2 = dz2.T @ a1, dz2.sum(0)
    dz1 = (dz2 @ W2) * gelu_grad(z1)
    gW1, gb1 = dz1.T @ x0, dz1.sum(0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def adam_step(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One textbook Adam update (Kingma & Ba 2015, Algorithm 1)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
