"""NumPy FP64 forward/backward of the 8 -> 64 -> 64 -> 1 SiLU proposal network
(TEST INFRASTRUCTURE ONLY; the network has no counterpart in the reference).

Parameter vector: W1[64][8] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3.  Loss: mean squared error.
Pinned by central finite differences of the loss (tests/test_oracle_mlp.py)."""

from __future__ import annotations

import numpy as np

H = 64
IN = 8


def unpack(p: np.ndarray):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    w1 = p[o:o + IN * H].reshape(H, IN); o += IN * H
    b1 = p[o:o + H]; o += H
    w2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return w1, b1, w2, b2, w3, b3


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(p: np.ndarray, x: np.ndarray) -> np.ndarray:
    w1, b1, w2, b2, w3, b3 = unpack(p)
    z1 = x @ w1.T + b1
    h1 = z1 * _sig(z1)
    z2 = h1 @ w2.T + b2
    h2 = z2 * _sig(z2)
    return h2 @ w3 + b3


def loss_and_gradient(p: np.ndarray, x: np.ndarray, t: np.ndarray):
    """Mean squared error and its gradient with respect to the parameter vector."""
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    w1, b1, w2, b2, w3, b3 = unpack(p)
    n = x.shape[0]
    z1 = x @ w1.T + b1
    s1 = _sig(z1)
    h1 = z1 * s1
    z2 = h1 @ w2.T + b2
    s2 = _sig(z2)
    h2 = z2 * s2
    y = h2 @ w3 + b3
    r = y - t
    d = 2.0 * r / n
    gw3 = h2.T @ d
    gb3 = d.sum()
    dz2 = (d[:, None] * w3[None, :]) * (s2 * (1.0 + z2 * (1.0 - s2)))
    gw2 = dz2.T @ h1
    gb2 = dz2.sum(axis=0)
    dh1 = dz2 @ w2
    dz1 = dh1 * (s1 * (1.0 + z1 * (1.0 - s1)))
    gw1 = dz1.T @ x
    gb1 = dz1.sum(axis=0)
    g = np.concatenate([gw1.ravel(), gb1, gw2.ravel(), gb2, gw3, [gb3]])
    return float(np.mean(r * r)), g
