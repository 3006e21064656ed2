"""FP64 numpy oracle of the warm-start MLP (132 -> 256 -> 256 -> 1, GELU).

TEST INFRASTRUCTURE ONLY: the checker for the GPU forward and backward
kernels (openmps_amd/csrc/mps_mlp.hip).  Parameter layout (as the library's
mps_model_get_params returns it):
  w1 [256][132], b1 [256], w2 [256][256], b2 [256], w3 [256], b3 [1]
  z1 = x w1^T + b1, h1 = gelu(z1), z2 = h1 w2^T + b2, h2 = gelu(z2),
  y = h2 . w3 + b3,  loss = mean((y - t)^2).
GELU is the tanh form 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))).
"""
from __future__ import annotations

import numpy as np

C0 = np.sqrt(2.0 / np.pi)
C1 = 0.044715


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C0 * (z + C1 * z ** 3)))


def gelu_grad(z):
    u = C0 * (z + C1 * z ** 3)
    th = np.tanh(u)
    return 0.5 * (1.0 + th) + 0.5 * z * (1.0 - th * th) * C0 * (1.0 + 3.0 * C1 * z * z)


def forward(p: dict, x):
    z1 = x @ p["w1"].T + p["b1"]
    h1 = gelu(z1)
    z2 = h1 @ p["w2"].T + p["b2"]
    h2 = gelu(z2)
    y = h2 @ p["w3"] + p["b3"][0]
    return y, (z1, h1, z2, h2)


def loss_and_grads(p: dict, x, t):
    """Mean squared error and its gradient with respect to every parameter."""
    m = x.shape[0]
    y, (z1, h1, z2, h2) = forward(p, x)
    d = y - t
    loss = float(np.mean(d * d))
    dy = 2.0 * d / m
    g = {"w3": h2.T @ dy, "b3": np.array([dy.sum()])}
    da2 = np.outer(dy, p["w3"]) * gelu_grad(z2)
    g["w2"] = da2.T @ h1
    g["b2"] = da2.sum(axis=0)
    dh1 = da2 @ p["w2"]
    da1 = dh1 * gelu_grad(z1)
    g["w1"] = da1.T @ x
    g["b1"] = da1.sum(axis=0)
    return loss, g


def to_bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    bits = a32.view(np.uint32).astype(np.uint64)
    rounded = ((bits + 0x7FFF + ((bits >> 16) & 1)) >> 16) << 16
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def numeric_grad(p: dict, x, t, name: str, idx, h: float = 1e-6) -> float:
    """Central difference of the loss in one parameter entry (pins the
    analytic gradient above)."""
    q = {k: v.copy() for k, v in p.items()}
    q[name][idx] += h
    lp, _ = loss_and_grads(q, x, t)
    q[name][idx] -= 2 * h
    lm, _ = loss_and_grads(q, x, t)
    return (lp - lm) / (2 * h)
