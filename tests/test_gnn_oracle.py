"""The FP64 restatement of the warm-start network: analytic gradients against
central finite differences, and the mean aggregation against its dense form."""

import numpy as np

from oracle import gnn_oracle as go


def _graph(n, rng):
    # random symmetric graph with one isolated node
    A = (rng.rand(n, n) < 0.3)
    A = np.triu(A, 1)
    A = A | A.T
    A[n - 1, :] = False
    A[:, n - 1] = False
    ptr = np.concatenate([[0], np.cumsum(A.sum(axis=1))]).astype(np.int64)
    idx = np.concatenate([np.nonzero(A[i])[0] for i in range(n)]).astype(np.int64)
    return A, ptr, idx


def test_mean_agg_dense_and_adjoint():
    rng = np.random.RandomState(3)
    A, ptr, idx = _graph(9, rng)
    h = rng.randn(9, 5)
    deg = np.maximum(A.sum(axis=1), 1)
    dense = (A / deg[:, None]) @ h
    assert np.allclose(go.mean_agg(h, ptr, idx), dense, rtol=0, atol=1e-14)
    g = rng.randn(9, 5)
    # <mean_agg(h), g> == <h, mean_agg_T(g)>
    lhs = np.sum(go.mean_agg(h, ptr, idx) * g)
    rhs = np.sum(h * go.mean_agg_T(g, ptr, idx))
    assert abs(lhs - rhs) < 1e-12


def test_gradients_match_finite_differences():
    rng = np.random.RandomState(5)
    n, fi, fo, H, L = 10, 3, 2, 6, 2
    _, ptr, idx = _graph(n, rng)
    X = rng.randn(n, fi)
    T = rng.randn(n, fo)
    mask = (rng.rand(n, fo) < 0.8).astype(np.float64)
    params = go.init_params(fi, fo, H, L, seed=1)
    for k in params:
        if k.startswith("b"):
            params[k] = rng.randn(*params[k].shape) * 0.1
    loss, g, _ = go.loss_and_grads(params, X, T, mask, ptr, idx, L)
    eps = 1e-6
    for name in go.param_names(L):
        P = params[name]
        flat = P.reshape(-1)
        for j in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[j]
            flat[j] = old + eps
            lp = go.loss_and_grads(params, X, T, mask, ptr, idx, L)[0]
            flat[j] = old - eps
            lm = go.loss_and_grads(params, X, T, mask, ptr, idx, L)[0]
            flat[j] = old
            fd = (lp - lm) / (2 * eps)
            an = g[name].reshape(-1)[j]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (name, j, fd, an)


def test_silu_grad_matches_finite_difference():
    z = np.linspace(-8, 8, 41)
    fd = (go.silu(z + 1e-6) - go.silu(z - 1e-6)) / 2e-6
    assert np.allclose(go.silu_grad(z), fd, atol=1e-8)
