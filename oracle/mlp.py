"""CPU restatement of the learned troubled-cell indicator (test infrastructure).

numpy, float64, following dg_amd/csrc/mlp.hip: the per-element features
(tensor-Legendre modes of the 2D p=2 two-fluid state, scaled, plus the
neighbour jumps of the cell means of rho, p and 1/(gamma-1)), the
49 -> 64 -> 64 -> 1 network with BF16 rounding of weights, inputs and
activations (round to nearest even, as the GPU), the weighted
cross-entropy and its gradient.  The gradient is pinned by finite
differences of the same loss (tests/test_mlp_oracle.py); the GPU gradient
kernels are checked against it.  The indicator has no reference
counterpart; its labels are the reference sensor's flags.
"""
from __future__ import annotations

import numpy as np

N_IN, N_PAD, N_H = 49, 64, 64
W1, B1 = 0, N_H * N_PAD
W2 = B1 + N_H
B2 = W2 + N_H * N_H
W3 = B2 + N_H
B3 = W3 + N_H
N_PARAM = B3 + 1


def bf16(x):
    """Round float32 values to BF16 (nearest even) and return them as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def legendre_vinv(ref_nodes: np.ndarray) -> np.ndarray:
    """Vinv[q, i] with q = a*3 + b: nodal values -> tensor-Legendre modes."""
    P = [lambda x: np.ones_like(x), lambda x: x, lambda x: 0.5 * (3 * x * x - 1)]
    xy = ref_nodes.reshape(-1, 2)
    V = np.stack([P[q // 3](xy[:, 0]) * P[q % 3](xy[:, 1]) for q in range(9)], axis=1)  # [i, q]
    return np.linalg.inv(V)


def features(U: np.ndarray, neighbors: np.ndarray, vinv: np.ndarray) -> np.ndarray:
    """float32 features [N_E, 64] of the 2D p=2 two-fluid state U[(e*5+f)*9+i]."""
    u = np.asarray(U, dtype=np.float64).reshape(-1, 5, 9)
    NE = u.shape[0]
    m = np.einsum("qi,efi->efq", vinv, u)  # modal coefficients
    mean = m[:, :, 0]
    gam = 1.0 + 1.0 / mean[:, 4]
    p = (gam - 1.0) * (mean[:, 3] - 0.5 * (mean[:, 1] ** 2 + mean[:, 2] ** 2) / mean[:, 0])
    avg = np.stack([mean[:, 0], p, mean[:, 4]], axis=1)
    sc_mom = np.sqrt(np.abs(mean[:, 0]) * np.abs(mean[:, 3]))
    scale = np.stack([np.abs(mean[:, 0]), sc_mom, sc_mom, np.abs(mean[:, 3]), np.abs(mean[:, 4])], axis=1)
    inv = np.where(scale > 0, 1.0 / np.where(scale > 0, scale, 1.0), 0.0)
    x = np.clip(m * inv[:, :, None], -8.0, 8.0).reshape(NE, 45)
    X = np.zeros((NE, N_PAD), dtype=np.float64)
    X[:, :45] = x
    nb = np.asarray(neighbors).reshape(NE, 4)
    for n in range(4):
        o = nb[:, n]
        ok = o >= 0
        oo = np.where(ok, o, 0)
        j = np.zeros(NE)
        for c in range(3):
            a, b = avg[oo, c], avg[:, c]
            s = np.abs(a) + np.abs(b)
            j = np.maximum(j, np.where(s > 0, np.abs(a - b) / np.where(s > 0, s, 1.0), 0.0))
        X[:, 45 + n] = np.where(ok, j, 0.0)
    return X.astype(np.float32)


def unpack(params):
    P = np.asarray(params, dtype=np.float64)
    return (P[W1:B1].reshape(N_H, N_PAD), P[B1:W2], P[W2:B2].reshape(N_H, N_H), P[B2:W3], P[W3:B3], P[B3])


def forward(params, X, round_bf16: bool = True):
    """Logits z and the cached activations (x, h1, h2) used by the gradient."""
    w1, b1, w2, b2, w3, b3 = unpack(params)
    r = bf16 if round_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    x = r(X)
    w1r, w2r, w3r = r(w1), r(w2), r(w3)
    h1 = r(np.maximum(x @ w1r.T + b1, 0.0))
    h2 = r(np.maximum(h1 @ w2r.T + b2, 0.0))
    z = h2 @ w3r + b3
    return z, (x, h1, h2, w2r, w3r)


def loss(params, X, y, pos_weight: float, round_bf16: bool = True) -> float:
    z, _ = forward(params, X, round_bf16)
    sp_neg = np.logaddexp(0.0, -z)
    sp_pos = np.logaddexp(0.0, z)
    return float(np.mean(pos_weight * y * sp_neg + (1.0 - y) * sp_pos))


def gradient(params, X, y, pos_weight: float, round_bf16: bool = True):
    """Mean gradient of the weighted cross-entropy (same layout as the parameters)."""
    z, (x, h1, h2, w2r, w3r) = forward(params, X, round_bf16)
    n = X.shape[0]
    sg = 1.0 / (1.0 + np.exp(-z))
    dz = (pos_weight * y * (sg - 1.0) + (1.0 - y) * sg) / n
    d2 = np.where(h2 > 0, dz[:, None] * w3r[None, :], 0.0)
    d1 = np.where(h1 > 0, d2 @ w2r, 0.0)
    g = np.zeros(N_PARAM)
    g[W1:B1] = (d1.T @ x).ravel()
    g[B1:W2] = d1.sum(0)
    g[W2:B2] = (d2.T @ h1).ravel()
    g[B2:W3] = d2.sum(0)
    g[W3:B3] = dz @ h2
    g[B3] = dz.sum()
    return g
