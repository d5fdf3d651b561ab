"""FP64 NumPy forward/backward of the warm-start graph network -- TEST INFRASTRUCTURE ONLY.

The network has no counterpart in the reference: it predicts the state the
plain init phase reaches (per-cell Na/K/Cl and Vmem) from per-cell geometry
and the cluster's parameters, on the cell-adjacency graph (``Cells.cell_nn``,
built like ``cells.py:1536-1564``). This module restates it in FP64 so that
tests can check the GPU kernels (BF16 MFMA, FP32 accumulation) against it; its
own analytic backward is checked against central finite differences.

Architecture (``n_layers`` message-passing layers, width ``H``)::

    h0 = silu(X W_in + b_in)
    for l in 1..L:  a = mean_{j in N(i)} h_j ;  h = h + silu([h, a] W_l + b_l)
    y  = h W_out + b_out
    loss = mean((y - t)^2 * mask)

``nbr_ptr``/``nbr_idx`` is the CSR of neighbours; the mean divides by
``max(deg, 1)``.
"""

from __future__ import annotations

from typing import Dict, List

import numpy as np


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def mean_agg(h, ptr, idx):
    deg = np.maximum(np.diff(ptr), 1).astype(np.float64)
    out = np.zeros_like(h)
    src = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
    np.add.at(out, src, h[idx])
    return out / deg[:, None]


def mean_agg_T(g, ptr, idx):
    """Adjoint of mean_agg: d h_j += sum_{i: j in N(i)} g_i / deg_i."""
    deg = np.maximum(np.diff(ptr), 1).astype(np.float64)
    gs = g / deg[:, None]
    src = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
    out = np.zeros_like(g)
    np.add.at(out, idx, gs[src])
    return out


def forward(params: Dict[str, np.ndarray], X, ptr, idx, n_layers: int):
    cache = {}
    z0 = X @ params["W_in"] + params["b_in"]
    h = silu(z0)
    cache["z0"], cache["h0"] = z0, h
    for l in range(n_layers):
        a = mean_agg(h, ptr, idx)
        cat = np.concatenate([h, a], axis=1)
        z = cat @ params[f"W{l}"] + params[f"b{l}"]
        cache[f"cat{l}"], cache[f"zl{l}"] = cat, z
        h = h + silu(z)
        cache[f"h{l + 1}"] = h
    y = h @ params["W_out"] + params["b_out"]
    return y, cache


def loss_and_grads(params, X, T, mask, ptr, idx, n_layers: int):
    y, cache = forward(params, X, ptr, idx, n_layers)
    n = mask.sum()
    diff = (y - T) * mask
    loss = float(np.sum(diff * diff) / n)
    dy = 2.0 * diff / n
    g = {}
    h = cache[f"h{n_layers}"]
    g["W_out"] = h.T @ dy
    g["b_out"] = dy.sum(axis=0)
    dh = dy @ params["W_out"].T
    H = h.shape[1]
    for l in reversed(range(n_layers)):
        dz = dh * silu_grad(cache[f"zl{l}"])
        g[f"W{l}"] = cache[f"cat{l}"].T @ dz
        g[f"b{l}"] = dz.sum(axis=0)
        dcat = dz @ params[f"W{l}"].T
        dh = dh + dcat[:, :H] + mean_agg_T(dcat[:, H:], ptr, idx)
    dz0 = dh * silu_grad(cache["z0"])
    g["W_in"] = X.T @ dz0
    g["b_in"] = dz0.sum(axis=0)
    return loss, g, y


def init_params(f_in: int, f_out: int, hidden: int, n_layers: int, seed: int = 0):
    rng = np.random.RandomState(seed)
    p = {"W_in": rng.randn(f_in, hidden) * np.sqrt(2.0 / f_in), "b_in": np.zeros(hidden)}
    for l in range(n_layers):
        p[f"W{l}"] = rng.randn(2 * hidden, hidden) * (0.5 / np.sqrt(2 * hidden))
        p[f"b{l}"] = np.zeros(hidden)
    p["W_out"] = rng.randn(hidden, f_out) * (1.0 / np.sqrt(hidden))
    p["b_out"] = np.zeros(f_out)
    return p


def param_names(n_layers: int) -> List[str]:
    names = ["W_in", "b_in"]
    for l in range(n_layers):
        names += [f"W{l}", f"b{l}"]
    return names + ["W_out", "b_out"]
