"""FP64 NumPy oracle of the start-vector network (test infrastructure only).

The network maps a bitstring (bits as +-1) through Linear -> GELU -> Linear ->
GELU -> Linear to (log-amplitude, sign logit). The reference has no such
network; this oracle pins the device kernels' forward pass, loss and
gradients. With ``rounding=None`` it is exact FP64 calculus, pinned by central
finite differences in ``tests/test_oracle.py``; with ``rounding=bf16_round``
it rounds to bfloat16 exactly where the device kernels do (the activations
and back-propagated signals that feed a matrix-core product), so the device
can be held to a tight tolerance. ``bf16_round`` is pinned against torch's
bfloat16 conversion.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

__all__ = ["encode", "forward", "loss_and_grads", "bf16_round", "adam", "gelu", "gelu_grad"]


def bf16_round(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def encode(keys, num_qubits, kpad, extra=None):
    """x[i, q] = 1 - 2 bit_q(key_i) for q < num_qubits, then the columns of
    ``extra`` (n x n_extra, used as given), 0 up to kpad."""
    keys = np.asarray(keys, dtype=np.uint64)
    bits = ((keys[:, None] >> np.arange(num_qubits, dtype=np.uint64)) & np.uint64(1)).astype(float)
    x = np.zeros((keys.size, kpad))
    x[:, :num_qubits] = 1.0 - 2.0 * bits
    if extra is not None:
        extra = np.asarray(extra, dtype=float).reshape(keys.size, -1)
        x[:, num_qubits:num_qubits + extra.shape[1]] = extra
    return x


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def _id(v):
    return v


def forward(params, x, rounding=None):
    """Returns y (n x 2) and the cache (z1, h1, z2, h2)."""
    r = rounding or _id
    w1, b1, w2, b2, w3, b3 = params
    z1 = x @ w1 + b1
    h1 = r(gelu(z1))
    z2 = h1 @ w2 + b2
    h2 = r(gelu(z2))
    y = h2 @ w3 + b3
    return y, (z1, h1, z2, h2)


def loss_and_grads(params, x, target, lam=1.0, rounding=None, wsum=None):
    """Weighted loss sum w[(y0-a)^2 + lam BCE(sigmoid(y1), p)] / wsum (default
    sum w) and its gradients with respect to (w1, b1, w2, b2, w3, b3). Returns
    (loss, grads, y, dy)."""
    r = rounding or _id
    w1, b1, w2, b2, w3, b3 = params
    a, p, w = target[:, 0], target[:, 1], target[:, 2]
    wsum = w.sum() if wsum is None else wsum
    y, (z1, h1, z2, h2) = forward(params, x, rounding)
    y0, y1 = y[:, 0], y[:, 1]
    bce = np.maximum(y1, 0) - y1 * p + np.log1p(np.exp(-np.abs(y1)))
    loss = float(np.sum(w * ((y0 - a) ** 2 + lam * bce)) / wsum)
    sg = 1.0 / (1.0 + np.exp(-y1))
    dy = np.stack([2.0 * w * (y0 - a) / wsum, lam * w * (sg - p) / wsum], axis=1)
    dyr = r(dy)
    gw3 = h2.T @ dyr
    gb3 = dy.sum(axis=0)
    dz2 = (dyr @ w3.T) * gelu_grad(z2)
    dz2r = r(dz2)
    gw2 = h1.T @ dz2r
    gb2 = dz2.sum(axis=0)
    dz1 = (dz2r @ w2.T) * gelu_grad(z1)
    gw1 = x.T @ r(dz1)
    gb1 = dz1.sum(axis=0)
    return loss, (gw1, gb1, gw2, gb2, gw3, gb3), y, dy


def adam(w, g, m, v, lr, b1, b2, eps, step):
    """One Adam step (Kingma & Ba 2015, Algorithm 1) in FP64; returns (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
