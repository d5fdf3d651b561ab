"""Warm-start graph network on the cell-adjacency graph.

Predicts, per cell, the state the init phase reaches (log-ratios of Na/K/Cl to
their initial values and Vmem) from per-cell geometry and the cluster's
parameters. Every arithmetic operation of the forward pass, the backward pass
and the Adam update runs in the library's own HIP kernels (BF16 MFMA products
with FP32 accumulation, graph aggregation, SiLU); PyTorch only allocates the
device buffers.

Architecture (``n_layers`` message-passing layers of width ``hidden``)::

    h0 = silu(X W_in + b_in)
    h_{l+1} = h_l + silu(h_l Wa_l + mean_nbr(h_l) Wb_l + b_l)
    y  = h_L W_out + b_out

``W_l`` is stored as the (2*hidden x hidden) stack of ``Wa_l`` over ``Wb_l``.
"""

from __future__ import annotations

import ctypes
from typing import Dict, List, Optional

import numpy as np

from . import _native

_vp = ctypes.c_void_p


def _p(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def init_params_np(f_in: int, f_out: int, hidden: int, n_layers: int, seed: int = 0) -> Dict[str, np.ndarray]:
    """He-style initial weights (seeded)."""
    rng = np.random.RandomState(seed)
    p = {"W_in": rng.randn(f_in, hidden) * np.sqrt(2.0 / f_in), "b_in": np.zeros(hidden)}
    for l in range(n_layers):
        p[f"W{l}"] = rng.randn(2 * hidden, hidden) * (0.5 / np.sqrt(2 * hidden))
        p[f"b{l}"] = np.zeros(hidden)
    p["W_out"] = rng.randn(hidden, f_out) * (1.0 / np.sqrt(hidden))
    p["b_out"] = np.zeros(f_out)
    return p


class WarmStartGNN:
    def __init__(self, f_in: int, f_out: int, hidden: int = 64, n_layers: int = 6, *,
                 device: Optional[str] = None, seed: int = 0, params: Optional[Dict[str, np.ndarray]] = None):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("no GPU visible: the network runs on the GPU only")
        self.torch = torch
        self.device = torch.device(device or "cuda")
        self.stream = torch.cuda.Stream(device=self.device)
        self.lib = _native.load()
        self.f_in, self.f_out, self.H, self.L = f_in, f_out, hidden, n_layers
        p = params if params is not None else init_params_np(f_in, f_out, hidden, n_layers, seed)
        with torch.cuda.device(self.device), torch.cuda.stream(self.stream):
            self.params = {k: torch.tensor(np.asarray(v, dtype=np.float32), device=self.device)
                           for k, v in p.items()}
            self.m = {k: torch.zeros_like(v) for k, v in self.params.items()}
            self.v = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.t = 0
        self._ws = None

    # ------------------------------------------------------------------
    def _s(self):
        return _vp(self.stream.cuda_stream)

    def _chk(self, rc: int, what: str) -> None:
        if rc != _native.BT_OK:
            raise RuntimeError(f"{what} failed ({rc})")

    def _empty(self, *shape):
        torch = self.torch
        with torch.cuda.device(self.device), torch.cuda.stream(self.stream):
            return torch.empty(shape, dtype=torch.float32, device=self.device)

    def _workspace(self, n: int):
        need = max(n * max(self.H, self.f_out), 64 * 2 * self.H * self.H, 1 << 16)
        if self._ws is None or self._ws.numel() < need:
            self._ws = self._empty(need)
        return self._ws

    def gemm(self, A, sam, sak, B, sbk, sbn, C, ldc, m, n, k, beta=0):
        ws = self._workspace(max(m, k))
        ksplit = 1
        if k > 4096:
            ksplit = min(64, (k + 4095) // 4096)
            while ksplit * m * n > ws.numel() and ksplit > 1:
                ksplit //= 2
        if beta and m * n * max(ksplit, 1) > ws.numel():
            raise RuntimeError("gemm workspace too small")
        self._chk(self.lib.bt_gnn_gemm(_p(A), sam, sak, _p(B), sbk, sbn, _p(C), ldc, m, n, k, beta,
                                       _p(ws), ksplit, self._s()), "bt_gnn_gemm")

    def colsum(self, G, n, f, out):
        ws = self._workspace(n)
        self._chk(self.lib.bt_gnn_colsum(_p(G), f, n, f, _p(out), _p(ws), ws.numel(), self._s()), "colsum")

    # ------------------------------------------------------------------
    def forward(self, X, ptr, idx, keep: bool = True):
        """X: (n, f_in) float32; ptr/idx: int32 CSR of the symmetric neighbour graph."""
        n, H = X.shape[0], self.H
        P = self.params
        z0 = self._empty(n, H)
        h = self._empty(n, H)
        self.gemm(X, self.f_in, 1, P["W_in"], H, 1, z0, H, n, H, self.f_in)
        self._chk(self.lib.bt_gnn_bias_silu(_p(z0), _p(P["b_in"]), _p(h), None, n, H, self._s()), "bias_silu")
        cache = {"X": X, "z0": z0, "h0": h}
        for l in range(self.L):
            agg = self._empty(n, H)
            self._chk(self.lib.bt_gnn_mean_agg(_p(h), H, _p(ptr), _p(idx), _p(agg), H, n, H, self._s()), "agg")
            z = self._empty(n, H)
            W = P[f"W{l}"]
            self.gemm(h, H, 1, W, H, 1, z, H, n, H, H)                       # h Wa
            Wb = W[H:]
            self.gemm(agg, H, 1, Wb, H, 1, z, H, n, H, H, beta=1)            # + agg Wb
            hn = self._empty(n, H)
            self._chk(self.lib.bt_gnn_bias_silu(_p(z), _p(P[f"b{l}"]), _p(hn), _p(h), n, H, self._s()), "bias_silu")
            cache[f"agg{l}"], cache[f"z{l + 1}"], cache[f"h{l + 1}"] = agg, z, hn
            h = hn
        y = self._empty(n, self.f_out)
        self.gemm(h, H, 1, P["W_out"], self.f_out, 1, y, self.f_out, n, self.f_out, H)
        self._chk(self.lib.bt_gnn_bias(_p(y), _p(P["b_out"]), n, self.f_out, self._s()), "bias")
        if keep:
            self._cache = cache
        return y

    def loss(self, Y, T, mask, n_valid: float):
        """Masked MSE over ``n_valid`` unmasked entries; returns (loss, dY)."""
        count = Y.numel()
        dY = self._empty(*Y.shape)
        parts = self._empty((count + 255) // 256)
        self._chk(self.lib.bt_gnn_mse(_p(Y), _p(T), _p(mask), count, ctypes.c_float(1.0 / max(n_valid, 1.0)),
                                      _p(dY), _p(parts), parts.numel(), self._s()), "mse")
        self.stream.synchronize()
        return float(parts.cpu().numpy().astype(np.float64).sum()), dY

    def backward(self, dY, ptr, idx) -> Dict[str, "object"]:
        c = self._cache
        n, H, fo = dY.shape[0], self.H, self.f_out
        P = self.params
        g = {k: self._empty(*v.shape) for k, v in P.items()}
        hL = c[f"h{self.L}"]
        self.gemm(hL, 1, H, dY, fo, 1, g["W_out"], fo, H, fo, n)            # h_L^T dY
        self.colsum(dY, n, fo, g["b_out"])
        dH = self._empty(n, H)
        self.gemm(dY, fo, 1, P["W_out"], 1, fo, dH, H, n, H, fo)            # dY W_out^T
        dz = self._empty(n, H)
        dagg = self._empty(n, H)
        for l in reversed(range(self.L)):
            self._chk(self.lib.bt_gnn_silu_bwd(_p(dH), _p(c[f"z{l + 1}"]), _p(dz), n * H, self._s()), "silu_bwd")
            W = P[f"W{l}"]
            gW = g[f"W{l}"]
            self.gemm(c[f"h{l}"], 1, H, dz, H, 1, gW[:H], H, H, H, n)        # h^T dz
            self.gemm(c[f"agg{l}"], 1, H, dz, H, 1, gW[H:], H, H, H, n)      # agg^T dz
            self.colsum(dz, n, H, g[f"b{l}"])
            self.gemm(dz, H, 1, W, 1, H, dH, H, n, H, H, beta=1)             # dH += dz Wa^T
            self.gemm(dz, H, 1, W[H:], 1, H, dagg, H, n, H, H)               # dz Wb^T
            self._chk(self.lib.bt_gnn_mean_agg_T(_p(dagg), H, _p(ptr), _p(idx), _p(dH), H, n, H, 1, self._s()),
                      "agg_T")
        self._chk(self.lib.bt_gnn_silu_bwd(_p(dH), _p(c["z0"]), _p(dz), n * H, self._s()), "silu_bwd")
        self.gemm(c["X"], 1, self.f_in, dz, H, 1, g["W_in"], H, self.f_in, H, n)
        self.colsum(dz, n, H, g["b_in"])
        return g

    def adam(self, grads, lr: float = 1e-3, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8) -> None:
        self.t += 1
        for k, p in self.params.items():
            self._chk(self.lib.bt_gnn_adam(_p(p), _p(grads[k]), _p(self.m[k]), _p(self.v[k]), p.numel(),
                                           ctypes.c_float(lr), ctypes.c_float(b1), ctypes.c_float(b2),
                                           ctypes.c_float(eps), self.t, self._s()), "adam")

    def params_np(self) -> Dict[str, np.ndarray]:
        self.stream.synchronize()
        return {k: v.cpu().numpy().astype(np.float64) for k, v in self.params.items()}

    def sync(self) -> None:
        self.stream.synchronize()
