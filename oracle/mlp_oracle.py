"""CPU oracle of the learned-sensor MLP (TEST INFRASTRUCTURE ONLY).

Forward and backward of 16 -> 32 -> 32 -> 1 (ReLU) with the BF16 rounding points of the GPU
kernels made explicit: inputs, W1, W2, H1 (input of layer 2), delta2 (input of W2^T delta2 and
of dW2), delta1 (input of dW1) are rounded to BF16 (round to nearest even); everything else is
FP32 arithmetic, evaluated here in float64 so the only gap to the GPU is FP32 accumulation
order.  Independently of those conventions the plain float64 gradient is pinned against central
finite differences in the tests.  The network has no counterpart in the reference.
"""
from __future__ import annotations

import numpy as np

NIN, NH = 16, 32


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float64."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_bits(a) -> np.ndarray:
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + NH * NIN].reshape(NH, NIN)
    o += NH * NIN
    b1 = p[o:o + NH]
    o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH)
    o += NH * NH
    b2 = p[o:o + NH]
    o += NH
    W3 = p[o:o + NH]
    o += NH
    return W1, b1, W2, b2, W3, p[o]


def forward(p, x, rounding=True):
    """x: (n, 16) float (already BF16 values when rounding); returns o (n,) and the activations."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    xr = r(x)
    z1 = xr @ r(W1).T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = r(h1) @ r(W2).T + b2
    h2 = np.maximum(z2, 0.0)
    o = h2 @ W3 + b3
    return o, dict(x=xr, z1=z1, h1=h1, z2=z2, h2=h2)


def grad(p, x, y, rounding=True):
    """Gradient of L = mean (o - y)^2 as one parameter vector, and L."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if rounding else (lambda a: np.asarray(a, dtype=np.float64))
    o, a = forward(p, x, rounding)
    n = x.shape[0]
    err = o - y
    loss = float(np.mean(err ** 2))
    dout = 2.0 * err / n
    gW3 = a["h2"].T @ dout
    gb3 = dout.sum()
    d2 = (a["z2"] > 0) * (dout[:, None] * W3[None, :])
    gb2 = d2.sum(0)
    d2r = r(d2)
    d1 = (a["z1"] > 0) * (d2r @ r(W2))
    gb1 = d1.sum(0)
    gW2 = d2r.T @ r(a["h1"])
    gW1 = r(d1).T @ a["x"]
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]]), loss


def finite_difference(p, x, y, idx, h=1e-6):
    """Central differences of the unrounded float64 loss for the parameters in idx."""
    out = []
    for i in idx:
        pp = np.array(p, dtype=np.float64)
        pm = pp.copy()
        pp[i] += h
        pm[i] -= h
        lp = np.mean((forward(pp, x, False)[0] - y) ** 2)
        lm = np.mean((forward(pm, x, False)[0] - y) ** 2)
        out.append((lp - lm) / (2 * h))
    return np.array(out)
