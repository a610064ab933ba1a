"""FP64 restatement of the learned eddy-viscosity network -- TEST
INFRASTRUCTURE ONLY (imported by tests/ to check the GPU kernels).

The network and loss are defined in wrf-jax_amd/csrc/k_closure.h:
    pre1 = W1 x + b1,  h1 = silu(pre1)         (x, W1 in bf16; h1 rounded to bf16)
    pre2 = W2 h1 + b2, h2 = silu(pre2)          (W2 in bf16)
    z    = w3 . h2 + b3
    K    = kref * softplus(z) / divisor
    loss = sum over rows of ((t1 + K s1)^2 + (t2 + K s2)^2) * lossnorm
Parameters are one vector theta (W1[64][64], b1, W2[64][64], b2, w3, b3).
``forward``/``gradient`` evaluate this in FP64 with the bf16 roundings of
the GPU path applied where the GPU applies them (``bf16=True``) or not at
all (``bf16=False``, the smooth function that the finite-difference check
pins); the gradient treats the roundings as the identity, as the GPU's
backward pass does.
"""

from __future__ import annotations

import numpy as np

NPARAMS = 8385
OFF = dict(W1=0, b1=4096, W2=4160, b2=8256, w3=8320, b3=8384)


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest, ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(theta: np.ndarray):
    t = np.asarray(theta, dtype=np.float64)
    W1 = t[OFF["W1"]:OFF["b1"]].reshape(64, 64)
    b1 = t[OFF["b1"]:OFF["W2"]]
    W2 = t[OFF["W2"]:OFF["b2"]].reshape(64, 64)
    b2 = t[OFF["b2"]:OFF["w3"]]
    w3 = t[OFF["w3"]:OFF["b3"]]
    b3 = t[OFF["b3"]]
    return W1, b1, W2, b2, w3, b3


def _sig(v):
    return 1.0 / (1.0 + np.exp(-v))


def _softplus(v):
    return np.where(v > 20.0, v, np.log1p(np.exp(np.minimum(v, 20.0))))


def forward(theta, X, bf16=True):
    """Network output z per row and the intermediates (FP64)."""
    W1, b1, W2, b2, w3, b3 = unpack(theta)
    X = np.asarray(X, dtype=np.float64)
    if bf16:
        X, W1, W2 = bf16_round(X), bf16_round(W1), bf16_round(W2)
    pre1 = X @ W1.T + b1
    h1 = pre1 * _sig(pre1)
    if bf16:
        h1 = bf16_round(h1)
    pre2 = h1 @ W2.T + b2
    h2 = pre2 * _sig(pre2)
    z = h2 @ w3 + b3
    return z, dict(X=X, W1=W1, W2=W2, w3=w3, pre1=pre1, h1=h1, pre2=pre2, h2=h2)


def loss_and_gradient(theta, X, T, lossnorm, kref, bf16=True):
    """Loss sum and its gradient with respect to theta (FP64)."""
    z, c = forward(theta, X, bf16)
    T = np.asarray(T, dtype=np.float64)
    t1, t2, s1, s2, div = T[:, 0], T[:, 1], T[:, 2], T[:, 3], T[:, 4]
    K = kref * _softplus(z) / div
    r1, r2 = t1 + K * s1, t2 + K * s2
    loss = float(np.sum((r1 * r1 + r2 * r2) * lossnorm))
    dK = 2.0 * (r1 * s1 + r2 * s2) * lossnorm
    dz = dK * kref * _sig(z) / div
    g = np.zeros(NPARAMS)
    g[OFF["b3"]] = dz.sum()
    g[OFF["w3"]:OFF["b3"]] = dz @ c["h2"]
    s2g = _sig(c["pre2"])
    d2 = dz[:, None] * c["w3"][None, :] * (s2g + c["pre2"] * s2g * (1.0 - s2g))
    g[OFF["b2"]:OFF["w3"]] = d2.sum(axis=0)
    g[OFF["W2"]:OFF["b2"]] = (d2.T @ c["h1"]).ravel()
    dh1 = d2 @ c["W2"]
    s1g = _sig(c["pre1"])
    d1 = dh1 * (s1g + c["pre1"] * s1g * (1.0 - s1g))
    g[OFF["b1"]:OFF["W2"]] = d1.sum(axis=0)
    g[OFF["W1"]:OFF["b1"]] = (d1.T @ c["X"]).ravel()
    return loss, g


def init_theta(seed: int = 0) -> np.ndarray:
    """He-style initial parameters (the GPU trainer starts from the same)."""
    rng = np.random.default_rng(seed)
    t = np.zeros(NPARAMS)
    t[OFF["W1"]:OFF["b1"]] = rng.standard_normal(4096) * np.sqrt(2.0 / 55.0)
    t[OFF["W2"]:OFF["b2"]] = rng.standard_normal(4096) * np.sqrt(2.0 / 64.0)
    t[OFF["w3"]:OFF["b3"]] = rng.standard_normal(64) * np.sqrt(1.0 / 64.0)
    return t
