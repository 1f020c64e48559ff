"""NumPy oracle of the rank-predictor MLP (32 -> 128 -> 128 -> 1, GELU).

TEST INFRASTRUCTURE ONLY.  The network has no counterpart in the reference
(its predictors, tools_qih/ai_rank_predictor.py:14-89 and
finetuned_predictor.py:8-23, act after a full SVD); the training recipe
follows ai_rank_predictor.py:117-195 (AdamW, MSE).  Parameter layout is the
flat vector  W1[128,32] | b1 | W2[128,128] | b2 | W3[128] | b3  used by the
HIP kernels.  Forward uses weights rounded to BF16 (round-to-nearest-even),
biases in full precision, FP64 arithmetic here.

Parity pinning: the backward pass is checked against central finite
differences of the forward (tests/test_mlp.py), AdamW against a hand-worked
single step.
"""
from __future__ import annotations

import math

import numpy as np

IN, H = 32, 128
O_W1 = 0
O_B1 = O_W1 + H * IN
O_W2 = O_B1 + H
O_B2 = O_W2 + H * H
O_W3 = O_B2 + H
O_B3 = O_W3 + H
NPARAM = O_B3 + 1

_erf = np.vectorize(math.erf)


def bf16_round(x: np.ndarray) -> np.ndarray:
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(p: np.ndarray, round_weights: bool = True):
    p = np.asarray(p, np.float64)
    w = bf16_round(p.astype(np.float32)).astype(np.float64) if round_weights else p
    W1 = w[O_W1:O_B1].reshape(H, IN)
    b1 = p[O_B1:O_W2]
    W2 = w[O_W2:O_B2].reshape(H, H)
    b2 = p[O_B2:O_W3]
    W3 = w[O_W3:O_B3]
    b3 = p[O_B3]
    return W1, b1, W2, b2, W3, b3


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


def forward(p, X, round_weights=True):
    W1, b1, W2, b2, W3, b3 = unpack(p, round_weights)
    z1 = X @ W1.T + b1
    h1 = gelu(z1)
    z2 = h1 @ W2.T + b2
    h2 = gelu(z2)
    return h2 @ W3 + b3, (z1, h1, z2, h2)


def backward(p, X, dy, round_weights=True):
    """Gradient of sum(dy * y) w.r.t. the flat parameter vector (weights are
    differentiated through as if the BF16 rounding were the identity)."""
    W1, b1, W2, b2, W3, b3 = unpack(p, round_weights)
    _, (z1, h1, z2, h2) = forward(p, X, round_weights)
    g = np.zeros(NPARAM)
    g[O_W3:O_B3] = dy @ h2
    g[O_B3] = dy.sum()
    d2 = (dy[:, None] * W3[None, :]) * gelu_grad(z2)
    g[O_W2:O_B2] = (d2.T @ h1).reshape(-1)
    g[O_B2:O_W3] = d2.sum(0)
    d1 = (d2 @ W2) * gelu_grad(z1)
    g[O_W1:O_B1] = (d1.T @ X).reshape(-1)
    g[O_B1:O_W2] = d1.sum(0)
    return g


def is_bias(idx: np.ndarray) -> np.ndarray:
    return ((idx >= O_B1) & (idx < O_W2)) | ((idx >= O_B2) & (idx < O_W3)) | (idx == O_B3)


def adamw(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=1e-5, scale=1.0):
    """One AdamW step (decoupled decay on weights only); returns (p, m, v)."""
    g = g * scale
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    decay = np.where(is_bias(np.arange(NPARAM)), 0.0, lr * wd)
    p = p - decay * p
    p = p - lr * mh / (np.sqrt(vh) + eps)
    return p, m, v


def init_params(seed: int) -> np.ndarray:
    """He-style init: W ~ N(0, 2/fan_in), biases 0."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM)
    p[O_W1:O_B1] = rng.normal(0, math.sqrt(2.0 / IN), H * IN)
    p[O_W2:O_B2] = rng.normal(0, math.sqrt(2.0 / H), H * H)
    p[O_W3:O_B3] = rng.normal(0, math.sqrt(1.0 / H), H)
    return p
